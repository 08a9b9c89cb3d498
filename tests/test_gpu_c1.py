"""The plain bulk Si run of the CPU reference path on the GPU.

10,000 electrons in the Si X valleys (Jacoboni phonons, m_t 0.19), 10 kV/cm
along [100], 300 K, 3000 steps of 1 fs, observables after every step
(examples/bulkSimulation/bulkSimulation.cpp:119-162, observables of
basicBulkParticleHandler.hpp:289-347).  The GPU (FP64 and FP32 kernels,
Philox streams) against the oracle's ensemble on the reference's own engine
(one mt19937_64 per particle): drift velocity, mean energy and the
occupancy of each of the three subvalley pairs, time-averaged over the last
2 ps, agree within 3 sigma.  Sigma is the standard error of each side from
20 equal particle groups (500 electrons each, independent streams), so it
includes the time correlation of the averages; the two sides' errors are
combined in quadrature.
"""
from concurrent.futures import ThreadPoolExecutor

import numpy as np
import pytest

from emc_models import (C1_AVG_FROM, C1_DT, C1_GROUPS, C1_PARTICLES, C1_STEPS, field_force,
                        group_stats, run_groups_gpu, si_bulk_c1)
from oracle import oracle_py as O

pytestmark = pytest.mark.gpu

NSIG = 3.0


def cpu_groups(spec, groups, per, steps, dt, seed, avg_from):
    """The oracle's reference-engine ensemble in `groups` independent runs,
    side by side on the host's cores.  Particle i of a run draws from the
    engine seeded with splitmix64(seed ^ i), so the runs' seeds differ above
    the particle bits (seed + g * 2^32)."""
    O.load()
    force, d = field_force(spec)
    Vt = O.thermal_voltage(spec["T"])
    deg = spec["valleys"][0]["deg"]

    def one(g):
        M = O.OracleModel(spec)
        obs, _, sub = M.run_mt(per, seed + (g << 32), Vt, steps, dt, force, d, threads=1,
                               sub_occ=True)
        return (obs[avg_from:, 1, 0].mean(), obs[avg_from:, 0, 0].mean(),
                sub[avg_from:, 0, :deg].mean(axis=0))

    with ThreadPoolExecutor(max_workers=8) as ex:
        res = list(ex.map(one, range(groups)))
    return dict(v=np.array([r[0] for r in res]), E=np.array([r[1] for r in res]),
                sub=np.array([r[2] for r in res]))


@pytest.fixture(scope="module")
def c1_cpu():
    spec = si_bulk_c1()
    print("C1 on the CPU reference path: %d groups x %d electrons x %d steps"
          % (C1_GROUPS, C1_PARTICLES // C1_GROUPS, C1_STEPS), flush=True)
    return cpu_groups(spec, C1_GROUPS, C1_PARTICLES // C1_GROUPS, C1_STEPS, C1_DT, 20261202,
                      C1_AVG_FROM)


@pytest.mark.parametrize("precision", [0, 1])
def test_c1_matches_cpu_reference_path(c1_cpu, precision):
    from viennaemc_amd import emcgpu
    spec = si_bulk_c1()
    print("C1 on the GPU, precision %d" % precision, flush=True)
    gpu = run_groups_gpu(emcgpu, spec, precision, C1_GROUPS, C1_PARTICLES // C1_GROUPS, C1_STEPS,
                         C1_DT, 20261202, C1_AVG_FROM)
    names = ["v_d [m/s]", "<E> [eV]", "occ sub 0", "occ sub 1", "occ sub 2"]
    g = np.column_stack([gpu["v"], gpu["E"], gpu["sub"]])
    c = np.column_stack([c1_cpu["v"], c1_cpu["E"], c1_cpu["sub"]])
    mg, sg = group_stats(g)
    mc, sc = group_stats(c)
    for i, name in enumerate(names):
        z = (mg[i] - mc[i]) / np.sqrt(sg[i] ** 2 + sc[i] ** 2)
        print("%-10s gpu %.6g +- %.2g   cpu %.6g +- %.2g   z = %+.2f"
              % (name, mg[i], sg[i], mc[i], sc[i], z), flush=True)
        assert abs(z) <= NSIG, name
    # physics sanity: electrons drift against the field at about 8e4 m/s
    assert mg[0] < 0 and 5e4 < -mg[0] < 1.2e5
