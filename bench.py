#!/usr/bin/env python3
"""Benchmark: particle-updates/s of the EMC particle step on one MI355X.

Workload (Si bulk, plain form): 10M electrons in the X valleys of Si
(m_l = 0.916, m_t = 0.19, alpha = 0.5/eV, three subvalley pairs), 300 K,
acoustic + six zero-order Jacoboni intervalley phonons (13 tables of 2000
levels on (0, 2] eV), 50 kV/cm along [100], dt = 1 fs.  A "step" advances
all particles by one dt including every scatter event inside it and the
per-step ensemble sums (drift velocity, mean energy, subvalley occupancy).

value      = FP32 (batched kernel, fused steps) particle-updates/s with the
             state resident in HBM; the FP64 plain kernel (one step per
             launch) is measured beside it.
roofline   = the FP32 step kernel: algorithmic bytes per launch / average
             launch time (HIP events on the library's stream).
cpu_baseline = the oracle's FP64 restatement of the reference loop with the
             reference engine (one mt19937_64 per particle), all host cores.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
for p in (ROOT, os.path.join(ROOT, "tests")):
    if p not in sys.path:
        sys.path.insert(0, p)

HBM_PEAK_GBS = 8000.0
BYTES_FP32 = 48  # k(3) + E + tau + valley|sub, 4 B each, read + write
BYTES_FP64 = 88  # k(3) + E + tau 8 B each + valley|sub 4 B, read + write


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=1000)
    ap.add_argument("--warmup", type=int, default=100)
    ap.add_argument("--particles", type=int, default=10_000_000)
    ap.add_argument("--spl", type=int, default=0, help="FP32 steps fused per launch (0 = auto)")
    ap.add_argument("--fp64-steps", type=int, default=0, help="FP64 timed steps (0 = --steps)")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-fp64", action="store_true")
    ap.add_argument("--no-plain", action="store_true", help="skip the FP32 one-step-per-launch leg")
    ap.add_argument("--ppl", type=int, default=0,
                    help="particles per GPU lane of the production kernel (0 = library default)")
    ap.add_argument("--no-batched", action="store_true", help="skip the 64-instance sweep leg")
    ap.add_argument("--batched-steps", type=int, default=2000)
    ap.add_argument("--no-hot", action="store_true", help="skip the hot-phonon cooling leg")
    ap.add_argument("--hot-steps", type=int, default=2000)
    ap.add_argument("--no-hot-plain", action="store_true",
                    help="run only the screened (q-resolved, plasma-screened) hot-phonon leg")
    ap.add_argument("--no-hot-screened", action="store_true",
                    help="skip the screened hot-phonon leg")
    ap.add_argument("--hot-particles", type=int, default=4_000_000)
    ap.add_argument("--no-device", action="store_true", help="skip the 2-D resistor leg")
    ap.add_argument("--device-only", action="store_true",
                    help="run only the 2-D resistor leg and print its record")
    ap.add_argument("--device-steps", type=int, default=20000)
    ap.add_argument("--device-particles", type=int, default=2_000_000)
    ap.add_argument("--device-fp64-steps", type=int, default=2000)
    ap.add_argument("--cpu-particles", type=int, default=1_000_000)
    ap.add_argument("--cpu-steps", type=int, default=300)
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="MIB",
                    help="only run a device copy of MIB MiB (rocprof byte-counter calibration)")
    return ap.parse_args()


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


def run_gpu(emcgpu, spec, precision, n, warmup, steps, spl, seed=20261202, ppl=0):
    ctx = emcgpu.Context(precision=precision)
    ctx.setup(0, spec)
    ctx.resize([n])
    ctx.init_particles(0, seed, spec["T"])
    if warmup:
        ctx.step(1e-15, warmup, seed=seed, step0=0, steps_per_launch=spl, ppl=ppl)
    ctx.synchronize()
    t0 = time.perf_counter()
    ctx.step(1e-15, steps, seed=seed, step0=warmup, steps_per_launch=spl, ppl=ppl)
    ctx.synchronize()
    wall = time.perf_counter() - t0
    av = ctx.averages(0)
    # timed pass for the kernel duration (HIP events around every launch)
    nt = max(spl, 1) * 4 if spl else 64
    ctx.timing(reset=True)
    ctx.step(1e-15, nt, seed=seed, step0=warmup + steps, steps_per_launch=spl, timing=True,
             ppl=ppl)
    ms, launches = ctx.timing()
    ctx.close()
    return dict(wall=wall, av=av, kernel_ms=ms / max(launches, 1), launches=launches,
                steps_per_launch=(nt / max(launches, 1)))


def run_batched(emcgpu, warmup, steps, seed=20261202):
    """The 64-instance Si bias/temperature sweep in one FP32 arena (tables
    built per instance in FP64, stored FP32), every call launched as one HIP
    graph with the launch shape (particles per lane = tile size / 256, steps
    fused per launch) chosen by emcgpu_autotune on this ensemble;
    whole-sweep particle-updates/s."""
    from emc_models import batched_instances
    inst = batched_instances(64, seed)
    counts = [n for _, n in inst]
    ctx = emcgpu.Context(precision=emcgpu.FP32, n_instances=len(inst))
    t0 = time.perf_counter()
    for i, (m, _) in enumerate(inst):
        ctx.setup(i, m)
    ctx.resize(counts)
    for i, (m, _) in enumerate(inst):
        ctx.init_particles(i, seed, m["T"])
    ctx.synchronize()
    setup_s = time.perf_counter() - t0
    # the tables of all instances again: per instance on the host, then in
    # one launch from the descriptors (the run uses the latter)
    nl, emax = inst[0][0]["n_levels"], inst[0][0]["max_energy"]
    mats = [dict(rho=m.get("rho", 2329.0), v_sound=m.get("v_sound", 9040.0)) for m, _ in inst]
    t0 = time.perf_counter()
    for i, (m, _) in enumerate(inst):
        ctx.build_tables(i, nl, emax, m["T"], **mats[i])
    tables_host_s = time.perf_counter() - t0
    t0 = time.perf_counter()
    ctx.build_tables_all(nl, emax, [m["T"] for m, _ in inst], mats)
    tables_gpu_s = time.perf_counter() - t0
    t0 = time.perf_counter()
    tune = ctx.autotune(1e-15, seed=seed, step0=0)
    tune_s = time.perf_counter() - t0
    if warmup:
        ctx.step(1e-15, warmup, seed=seed, step0=0, graph=True)
    ctx.synchronize()
    t0 = time.perf_counter()
    ctx.step(1e-15, steps, seed=seed, step0=warmup, graph=True)
    ctx.synchronize()
    wall = time.perf_counter() - t0
    av = [ctx.averages(i) for i in range(len(inst))]
    ctx.close()
    total = sum(counts)
    vd = [float(a["avgV"][-1, 0]) for a in av]
    return {"updates_per_s": total * steps / wall, "instances": len(inst), "particles": total,
            "steps": steps, "ms_per_step": wall / steps * 1e3, "setup_s": setup_s,
            "tables_host_s": tables_host_s, "tables_gpu_s": tables_gpu_s,
            "vd_min": min(vd), "vd_max": max(vd),
            "autotune": {"particles_per_lane": tune["best"][0], "steps_per_launch": tune["best"][1],
                         "tune_s": tune_s,
                         "trials_ms_per_step": [[p, k, round(ms, 5)] for p, k, ms in tune["trials"]]},
            "workload": "64 Si bulk instances, field log-U[0.1,200] kV/cm along [100]/[111], "
                        "T U[77,500] K, N log-U[1e4,2e6], FP32, HIP-graph launch, autotuned shape"}


def run_hot_phonon(emcgpu, n, warmup, steps, precision, seed=20261202, screened=False):
    """Photoexcited cooling with a hot-phonon bath: every step is the
    particle step with LDS bath tallies, the bath update and the table
    rebuild (one HIP graph per call); whole-job particle-updates/s.
    screened: the hot-carrier default of the reference (screened hot-phonon
    Froehlich, q-resolved rate and angle, plasma screening refreshed after
    every step from the step's tile partials)."""
    from emc_models import photoexcited_cooling, screened_cooling
    spec = screened_cooling(n) if screened else photoexcited_cooling(n)
    ctx = emcgpu.Context(precision=precision)
    ctx.setup(0, spec)
    ctx.resize([n])
    ctx.init_particles_energy(0, seed, 0.5, 0.05)
    if screened:
        ctx.update_screening(0)
    dt = 0.5e-15
    if warmup:
        ctx.step(dt, warmup, seed=seed, step0=0, graph=True)
    ctx.synchronize()
    t0 = time.perf_counter()
    ctx.step(dt, steps, seed=seed, step0=warmup, graph=True)
    ctx.synchronize()
    wall = time.perf_counter() - t0
    av = ctx.averages(0)
    hist = ctx.bath_history(0, 0)
    bath = ctx.get_bath(0, 0)
    qs2 = ctx.get_screening(0)
    ctx.timing(reset=True)
    ctx.step(dt, 64, seed=seed, step0=warmup + steps, timing=True)
    ms, launches = ctx.timing()
    ctx.close()
    if screened:
        return {"updates_per_s": n * steps / wall, "particles": n, "steps": steps,
                "ms_per_step": wall / steps * 1e3, "step_kernel_ms": ms / max(launches, 1),
                "meanE_first": float(av["avgE"][0, 0]), "meanE_last": float(av["avgE"][-1, 0]),
                "meanNq_last": float(hist[-1]), "qs2_last": float(qs2),
                "workload": "screened photoexcited cooling, %d e- N(0.5, 0.05) eV, screened hot-phonon "
                            "Froehlich with q-resolved rate and angle on the 256-bin bath, plasma "
                            "screening (eps 3.5) and window sums refreshed every step, dt 0.5 fs, "
                            "%s" % (n, "FP32" if precision else "FP64")}
    return {"updates_per_s": n * steps / wall, "particles": n, "steps": steps,
            "ms_per_step": wall / steps * 1e3, "step_kernel_ms": ms / max(launches, 1),
            "meanE_first": float(av["avgE"][0, 0]), "meanE_last": float(av["avgE"][-1, 0]),
            "meanNq_last": float(hist[-1]), "maxNq_last": float(bath["nq"].max()),
            "workload": "photoexcited cooling, %d e- N(0.5, 0.05) eV, Kane Gamma valley, hot-phonon "
                        "Froehlich on a 256-bin LO bath (tau_LO 1 ps into the acoustic reservoir, "
                        "30 ps to the lattice), dt 0.5 fs, bath update + table rebuild every step, "
                        "%s" % (n, "FP32" if precision else "FP64")}


def run_device(emcgpu, n, warmup, steps, precision, seed=20261202, graph=True):
    """Self-consistent 2-D resistor: 200 x 50 nodes at 1 nm, N_D 1e17 cm^-3
    with 1e19 cm^-3 ohmic regions, 0.5 V, parabolic m* = 0.26 valley with
    acoustic + intervalley phonons, CIC, Poisson (Newton + FP32 multigrid,
    FP64 residual to 1e-10) every 0.1 fs step; whole-job particle-updates/s
    (particles moved per step x steps / wall time, HIP-graph launch)."""
    from emc_models import MODELS_DEVICE, resistor_device
    dev = resistor_device(n_particles=n)
    ctx = emcgpu.Context(precision=precision)
    ctx.setup(0, MODELS_DEVICE["device_para"])
    ctx.device_setup(dev, pm=emcgpu.PM_CIC, poisson_tol=1e-10, poisson_max_iter=50)
    t0 = time.perf_counter()
    n0 = ctx.device_init(seed)
    t_init = time.perf_counter() - t0
    dt = 0.1e-15
    if warmup:
        ctx.device_step(dt, warmup, seed=seed, graph=graph)
    ctx.synchronize()
    t0 = time.perf_counter()
    out = ctx.device_step(dt, steps, seed=seed, graph=graph)
    ctx.synchronize()
    wall = time.perf_counter() - t0
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
            "workload": "2-D n+ n n+ resistor, 200 x 50 nodes at 1 nm, N_D 1e17 / 1e19 cm^-3, "
                        "0.5 V, %d e- (parabolic 0.26 m0, acoustic + IV), CIC, Poisson every "
                        "0.1 fs step, %s" % (n, "FP32" if precision else "FP64")}


PROFILE = os.path.join(ROOT, "profiles", "emc_step_si_bulk_10M.md")


def measured_traffic(n, spl):
    """Calibrated HBM bytes per FP32 launch from the committed rocprof counter
    summary (FETCH_SIZE / WRITE_SIZE scaled by a copy kernel), when this run
    has the profiled shape (10M particles, 64 steps per launch); else None."""
    if n != 10_000_000 or spl != 64 or not os.path.exists(PROFILE):
        return None
    for line in open(PROFILE):
        if line.startswith("| FP32 launch HBM traffic, calibrated"):
            return float(line.split("|")[2].split()[0]) * 1e6
    return None


def cpu_baseline(spec, n, steps):
    from oracle import oracle_py as O
    from emc_models import field_force
    M = O.OracleModel(spec)
    force, d = field_force(spec)
    threads = min(os.cpu_count() or 1, 16)
    t = np.zeros(1)
    M.run_mt(n, 20261202, O.thermal_voltage(spec["T"]), steps, 1e-15, force, d, obs=False,
             threads=threads, timing=t)
    return dict(value=n * steps / t[0], unit="particle-updates/s", cores=threads, kind="port",
                sample="%d particles x %d steps, FP64, mt19937_64 per particle (step loop only, %.1f s)"
                       % (n, steps, t[0]))


def main():
    args = parse()
    if args.copy_calibration:
        copy_calibration(args.copy_calibration)
        return 0
    if args.gpus != 1:
        print("bench.py measures one GPU: the particle step has no multi-GPU path in this library "
              "(--gpus %d requested)" % args.gpus, file=sys.stderr)
        return 2
    from emc_models import MODELS
    from viennaemc_amd import emcgpu
    spec = MODELS["si_jacoboni"]
    n = args.particles
    if args.device_only:
        rec = run_device(emcgpu, args.device_particles, args.warmup, args.device_steps,
                         emcgpu.FP32)
        print(json.dumps({"device_resistor": rec}), flush=True)
        return 0

    r32 = run_gpu(emcgpu, spec, emcgpu.FP32, n, args.warmup, args.steps, args.spl,
                  ppl=args.ppl)
    ups32 = n * args.steps / r32["wall"]
    out = {}
    if not args.no_fp64:
        s64 = args.fp64_steps or args.steps
        # the plain FP64 path: one particle per lane, one step per launch
        r64 = run_gpu(emcgpu, spec, emcgpu.FP64, n, args.warmup, s64, 1, ppl=1)
        ups64 = n * s64 / r64["wall"]
        # FP64 with the batched kernel's launch shape, for reference
        r64b = run_gpu(emcgpu, spec, emcgpu.FP64, n, args.warmup, s64, args.spl, ppl=args.ppl)
        out["fp64_batched_updates_per_s"] = n * s64 / r64b["wall"]
        tail = max(1, min(args.steps, s64) // 2)
        v32, v64 = r32["av"]["avgV"][-tail:, 0].mean(), r64["av"]["avgV"][-tail:, 0].mean()
        e32, e64 = r32["av"]["avgE"][-tail:, 0].mean(), r64["av"]["avgE"][-tail:, 0].mean()
        out.update(fp64_plain_updates_per_s=ups64, fp32_over_fp64_plain=ups32 / ups64,
                   fp64_kernel_ms=r64["kernel_ms"],
                   fp64_roofline_frac=(BYTES_FP64 * n / (r64["kernel_ms"] * 1e-3)) / 1e9 / HBM_PEAK_GBS,
                   delta_vd_rel=(v32 - v64) / v64, delta_E_rel=(e32 - e64) / e64,
                   vd_fp32=v32, vd_fp64=v64, meanE_fp32=e32, meanE_fp64=e64)
    if not args.no_plain:
        r1 = run_gpu(emcgpu, spec, emcgpu.FP32, n, 0, min(args.steps, 200), 1, ppl=args.ppl)
        out["fp32_plain_updates_per_s"] = n * min(args.steps, 200) / r1["wall"]
        out["fp32_plain_kernel_ms"] = r1["kernel_ms"]
        out["fp32_plain_roofline_frac"] = BYTES_FP32 * n / (r1["kernel_ms"] * 1e-3) / 1e9 / HBM_PEAK_GBS
    if not args.no_batched:
        out["batched_sweep"] = run_batched(emcgpu, args.warmup, args.batched_steps)
    if not args.no_hot:
        if not args.no_hot_plain:
            out["hot_phonon"] = run_hot_phonon(emcgpu, args.hot_particles, args.warmup,
                                               args.hot_steps, emcgpu.FP32)
            if not args.no_fp64:
                out["hot_phonon_fp64"] = run_hot_phonon(emcgpu, args.hot_particles, args.warmup,
                                                        args.hot_steps, emcgpu.FP64)
        if not args.no_hot_screened:
            out["hot_phonon_screened"] = run_hot_phonon(emcgpu, args.hot_particles, args.warmup,
                                                        args.hot_steps, emcgpu.FP32, screened=True)
    if not args.no_device:
        out["device_resistor"] = run_device(emcgpu, args.device_particles, args.warmup,
                                            args.device_steps, emcgpu.FP32)
        if not args.no_fp64:
            out["device_resistor_fp64"] = run_device(emcgpu, args.device_particles, args.warmup,
                                                     args.device_fp64_steps, emcgpu.FP64)
    spl = r32["steps_per_launch"]
    # algorithmic bytes per launch: 48 B per particle-update (the SoA state
    # read and written once per update) x the updates one launch processes;
    # with fused steps the HBM traffic itself is 1/spl of this
    achieved = BYTES_FP32 * n * spl / (r32["kernel_ms"] * 1e-3) / 1e9
    res = {
        "metric": "particle-updates/s (Si bulk, 10M e-) FP32-batched vs FP64 plain; % HBM peak; dv_d",
        "value": ups32,
        "unit": "particle-updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": r32["wall"] / args.steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (thermal Maxwellian init, Philox-4x32-10)",
        "config": {"workload": "Si bulk X-valley EMC, 10M e-, 50 kV/cm [100], dt 1 fs, 300 K, "
                               "acoustic + 6 zero-order Jacoboni phonons, 2000-level tables",
                   "particles": n, "steps_per_launch": spl, "particles_per_lane": args.ppl},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": measured_traffic(n, spl)},
    }
    res.update(out)
    if not args.no_cpu:
        res["cpu_baseline"] = cpu_baseline(spec, args.cpu_particles, args.cpu_steps)
    print(json.dumps(res), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
