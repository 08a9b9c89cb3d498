"""CPU calibration on the plain bulk Si run (10k electrons, 10 kV/cm [100],
300 K, 3000 steps of 1 fs): the real basicBulkParticleHandler loop
(oracle/_ref/ref_harness --time-c1, built from the ViennaEMC sources) against
the oracle's restatement (oracle_py.OracleModel.run_mt), both on one thread
with the per-step observables, timed in the same process environment.  The
ratio scales the port-based CPU baselines of bench.py to the reference's own
code.  Writes the summary (markdown with `| figure: ... |` rows) to stdout.

usage: python oracle/calibrate_c1.py [--particles N] [--steps S] > profiles/cpu_calibration_c1.md
"""
import argparse
import json
import os
import platform
import subprocess
import sys
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--particles", type=int, default=10_000)
    ap.add_argument("--steps", type=int, default=3000)
    ap.add_argument("--seed", type=int, default=20261202)
    ap.add_argument("--repeats", type=int, default=3)
    a = ap.parse_args()
    from oracle import oracle_py as O
    from viennaemc_amd.workloads import field_force, si_bulk_c1
    harness = os.path.join(HERE, "_ref", "ref_harness")
    if not os.path.exists(harness):
        print("build the harness first: make -C oracle ref", file=sys.stderr)
        return 2
    ref = []
    # the reference writes its scatter-table files into the working folder
    work = tempfile.mkdtemp(prefix="emc_c1_")
    for r in range(a.repeats):
        out = subprocess.check_output([harness, "--time-c1", str(a.particles), str(a.steps),
                                       str(a.seed + r)], text=True, cwd=work)
        ref.append(json.loads([l for l in out.splitlines() if l.startswith("{")][-1]))
    spec = si_bulk_c1()
    force, d = field_force(spec)
    Vt = O.thermal_voltage(spec["T"])
    port = []
    for r in range(a.repeats):
        t = np.zeros(1)
        n = ref[r]["particles"]  # same ensemble size as the handler created
        O.OracleModel(spec).run_mt(n, a.seed + r, Vt, a.steps, 1e-15, force, d, obs=True,
                                   threads=1, timing=t)
        port.append(n * a.steps / t[0])
    ups_ref = max(x["updates_per_s"] for x in ref)
    ups_port = max(port)
    cpu = platform.processor() or platform.machine()
    try:
        with open("/proc/cpuinfo") as f:
            names = [l.split(":", 1)[1].strip() for l in f if l.startswith("model name")]
        cpu = names[0] if names else cpu
    except OSError:
        pass
    print("# CPU calibration: reference handler loop vs the oracle port (plain bulk Si, C1 inputs)\n")
    print("Command (repository root, after `make -C oracle ref` where the ViennaEMC sources are):\n")
    print("```\npython oracle/calibrate_c1.py --particles %d --steps %d > profiles/cpu_calibration_c1.md\n```\n"
          % (a.particles, a.steps))
    print("Workload: Si X valleys (m_l 0.916, m_t 0.19, alpha 0.5/eV), acoustic + six zero-order "
          "Jacoboni phonons, 2000 levels to 2 eV, 10 kV/cm along [100], 300 K, dt 1 fs, %d steps, "
          "~%d electrons, the three per-step observables of bulkSimulation.cpp:148-157 after every "
          "step.  Reference: `basicBulkParticleHandler::moveParticles` compiled with g++ -O2 from the "
          "ViennaEMC sources (one shared mt19937_64 per thread, no OpenMP in this build).  Port: "
          "`orc_run_mt` of oracle/emc_oracle.c (one mt19937_64 per particle), gcc -O2.  Both on one "
          "thread of %s; best of %d runs each.\n" % (a.steps, a.particles, cpu, a.repeats))
    print("| run | particles | seconds | particle-updates/s |")
    print("|---|---|---|---|")
    for r, x in enumerate(ref):
        print("| reference %d | %d | %.3f | %.4g |" % (r, x["particles"], x["seconds"],
                                                   x["updates_per_s"]))
    for r, u in enumerate(port):
        print("| port %d | %d | %.3f | %.4g |" % (r, ref[r]["particles"],
                                                ref[r]["particles"] * a.steps / u, u))
    print("\n| figure: reference_updates_per_s_1thread | %.6g | |" % ups_ref)
    print("|---|---|---|")
    print("| figure: port_updates_per_s_1thread | %.6g | |" % ups_port)
    print("| figure: reference_over_port | %.4f | reference-equivalent CPU rate = port rate x this |"
          % (ups_ref / ups_port))
    return 0


if __name__ == "__main__":
    sys.exit(main())
