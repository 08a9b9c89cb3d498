"""Screened hot-phonon Froehlich with plasma screening on the GPU (the
reference's hot-carrier default: emcScreenedHotPhononFroehlich* with the
q-resolved rate and angle, emcPlasmonScreening refreshed after every step).

* Replay (FP64) against the reference's own run (golden fixtures recorded
  through emcScreenedHotPhononFroehlich*, emcPhononBath and the
  hotCarrierMHP screening update): scatter events per particle and step
  exactly, the bath's mean occupation and q_s^2 after every fifth step, and
  the rebuilt tables against the oracle's for the final bath and q_s^2.
  q_s^2 is taken from the GPU's per-tile energy sums, whose order differs
  from the reference's particle-order sum, and the screened log factor from
  the device's log: both agree to ~1e-16, hence rtol 1e-12 on q_s^2.
* Production (FP64) against the oracle's Philox twin: particle state, the
  bath and q_s^2 after every step.
* FP32 against FP64 on the cooling workload, and graph against eager launch.
"""
import numpy as np
import pytest

from emc_models import MODELS, field_force, screened_cooling
from oracle import oracle_py as O

pytestmark = pytest.mark.gpu

REL_TOL = 1e-12


def _emc():
    from viennaemc_amd import emcgpu
    return emcgpu


def _rel_k(a, b):
    return np.linalg.norm(a - b, axis=-1) / np.maximum(np.linalg.norm(b, axis=-1), 1e-300)


def _debye_qs2(spec, E, valley):
    """hotCarrierMHP's updateScreening for one species, summed in particle
    order."""
    Q, EPS0, KB = 1.60219e-19, 8.85419e-12, 1.38066e-23
    scr = spec["screening"]
    e0 = E[valley == 0]
    avg = sum(float(x) for x in e0) / len(e0) if len(e0) else 0.0
    ts = 2.0 * avg * Q / (3.0 * KB)
    if len(E) == 0 or ts <= 0:
        return 0.0
    return len(E) / scr["volume"] * Q * Q / (scr["eps"] * EPS0 * KB * ts)


def _ctx(spec, n, precision):
    emcgpu = _emc()
    ctx = emcgpu.Context(precision=precision)
    ctx.setup(0, spec)
    ctx.resize([n])
    return ctx


def test_replay_screened_hot_matches_reference(golden):
    emcgpu = _emc()
    name = "screened_hot_polar"
    spec = MODELS[name]
    b = name + ".traj."
    np_, ns, seed = [int(x) for x in golden[b + "meta"]]
    dt = golden[b + "dt"][0]
    print("screened hot-phonon replay vs reference:", np_, "particles", ns, "steps", flush=True)
    draws = O.mt64_particle_draws(seed, np_, 2048)
    cursor = golden[b + "init_draws"].astype(np.uint64).copy()
    grain = golden[b + "init_grain"].copy()
    ctx = _ctx(spec, np_, emcgpu.FP64)
    ctx.set_particles(0, golden[b + "init_k"], golden[b + "init_E"], golden[b + "init_tau"],
                      golden[b + "init_valley"], golden[b + "init_sub"])
    ctx.update_screening(0)
    mean, qs2, events, own, energies = [], [], [], [], []
    for s0 in range(0, ns, 5):
        tr = ctx.step(dt, 5, mode=emcgpu.REPLAY, step0=s0, draws=draws, cursor=cursor,
                      grain_tau=grain, trace=True)
        mean.append(ctx.bath_history(0, 0))
        qs2.append(ctx.get_screening(0))
        events.append(tr["events"])
        p = ctx.get_particles(0)
        energies.append(p["E"])
        own.append(_debye_qs2(spec, p["E"], p["valley"]))
    mean = np.concatenate(mean)
    qs2, own = np.array(qs2), np.array(own)
    assert (np.concatenate(events) == golden[b + "nevents"].reshape(ns, np_)).all()
    assert np.allclose(mean, golden[b + "bath_mean"], rtol=1e-13, atol=0)
    # q_s^2 is the Debye form of the particles' own energies (the GPU's
    # tile sums against a sequential sum)
    assert np.allclose(qs2, own, rtol=1e-14, atol=0)
    ref_qs2 = golden[b + "qs2"][4::5]
    Eref = golden[b + "E"].reshape(-1, np_)
    dE = np.abs(np.array(energies) - Eref).max(axis=1) / Eref.mean(axis=1)
    print("q_s^2 %.4e -> %.4e, max rel diff %.2e; max |dE| / <E> %.2e" % (
        ref_qs2[0], ref_qs2[-1], (np.abs(qs2 - ref_qs2) / ref_qs2).max(), dE.max()), flush=True)
    # and the reference's within the trajectories' own agreement
    assert np.allclose(qs2, ref_qs2, rtol=max(1e-12, 64 * dE.max()), atol=0)
    assert ref_qs2[-1] > 2 * ref_qs2[0]  # the carriers cooled: stronger screening
    # the rebuilt device tables are the reference's for that bath and q_s^2
    M = O.OracleModel(spec)
    M.set_bath(0, ctx.get_bath(0, 0)["nq"])
    M.set_qs2(qs2[-1])
    M.L.orc_model_build_tables(M.h)
    cum, tau0 = ctx.get_tables(0, 0, spec["n_levels"])
    assert abs(tau0 - M.tau0(0)) <= 1e-12 * M.tau0(0)
    assert np.allclose(cum, M.cum(0), rtol=1e-12, atol=1e-16)


def _twin(spec, n, ns, seed, sigma=0.05):
    emcgpu = _emc()
    M = O.OracleModel(spec)
    ref = M.init_philox_energy(n, spec["init_energy"], sigma, seed)
    ctx = _ctx(spec, n, emcgpu.FP64)
    ctx.set_particles(0, ref["k"], ref["E"], ref["tau"], ref["valley"], ref["sub"])
    M.seed_screening(ref["E"], ref["valley"])
    ctx.update_screening(0)
    assert abs(ctx.get_screening(0) - M.qs2()) <= 1e-13 * M.qs2()
    force, d = field_force(spec)
    ctx.step(0.5e-15, ns, seed=seed)
    hist, qs2 = [], []
    for s in range(ns):
        M.run_philox(ref, 1, 0.5e-15, force, d, seed, step0=s, obs=False)
        hist.append(M.bath(0)["mean"])
        qs2.append(M.qs2())
    return ctx, M, ref, np.array(hist), np.array(qs2)


def test_production_screened_hot_matches_oracle_twin():
    """The discrete history exactly; k and E within the spread that one ulp
    of the initial k causes in the oracle itself (the q-resolved angle turns
    rounding at the window edge into direction noise of ~sqrt(ulp)), and at
    most 1e-6 / 1e-8; the bath and q_s^2 within 1e-10."""
    spec = MODELS["screened_hot_polar"]
    n, ns, seed = 4096, 60, 0x5C4EE
    print("screened hot-phonon production vs twin", flush=True)
    ctx, M, ref, hist, qs2 = _twin(spec, n, ns, seed)
    # the oracle's own sensitivity on the same run
    M2 = O.OracleModel(spec)
    ref2 = M2.init_philox_energy(n, spec["init_energy"], 0.05, seed)
    ref2["k"] *= 1.0 + 2.0 ** -52
    M2.seed_screening(ref2["E"], ref2["valley"])
    force, d = field_force(spec)
    for s in range(ns):
        M2.run_philox(ref2, 1, 0.5e-15, force, d, seed, step0=s, obs=False)
    sk = _rel_k(ref2["k"].reshape(-1, 3), ref["k"].reshape(-1, 3)).max()
    se = (np.abs(ref2["E"] - ref["E"]) / ref["E"]).max()
    g = ctx.get_particles(0)
    dk = _rel_k(g["k"].reshape(-1, 3), ref["k"].reshape(-1, 3)).max()
    de = (np.abs(g["E"] - ref["E"]) / ref["E"]).max()
    print("GPU vs twin: k %.2e E %.2e; 1-ulp sensitivity k %.2e E %.2e" % (dk, de, sk, se),
          flush=True)
    assert (g["valley"] == ref["valley"]).all() and (g["sub"] == ref["sub"]).all()
    assert dk < max(REL_TOL, min(10 * sk, 1e-6))
    assert de < max(REL_TOL, min(10 * se, 1e-8))
    assert np.allclose(ctx.bath_history(0, 0), hist, rtol=1e-10, atol=0)
    assert np.allclose(ctx.get_bath(0, 0)["nq"], M.bath(0)["nq"], rtol=1e-10, atol=0)
    assert abs(ctx.get_screening(0) - qs2[-1]) <= 1e-10 * qs2[-1]
    assert qs2[-1] != qs2[0]


@pytest.mark.parametrize("precision", [0, 1])
def test_graph_launch_with_screening_matches_eager(precision):
    emcgpu = _emc()
    n = 50_000
    spec = screened_cooling(n)
    print("screened graph vs eager, precision", precision, flush=True)
    out = {}
    for graph in (False, True):
        ctx = _ctx(spec, n, precision)
        ctx.init_particles_energy(0, 37, 0.5, 0.05)
        ctx.update_screening(0)
        rec = []
        step0 = 0
        for ns in (30, 20, 45):
            ctx.step(0.5e-15, ns, seed=37, step0=step0, graph=graph)
            step0 += ns
            p = ctx.get_particles(0)
            rec.append((p["k"], p["E"], ctx.get_bath(0, 0)["nq"], ctx.bath_history(0, 0),
                        np.array([ctx.get_screening(0)])))
        ctx.close()
        out[graph] = rec
    for a, b in zip(out[False], out[True]):
        for x, y in zip(a, b):
            assert np.array_equal(x, y)


def test_fp32_screened_cooling_matches_fp64():
    """Screened, q-resolved photoexcited cooling for 0.3 ps: FP32 and FP64
    mean energy, phonon occupation and q_s^2 agree; screening slows the
    cooling against the unscreened bath."""
    emcgpu = _emc()
    n, ns = 400_000, 600
    spec = screened_cooling(n)
    res = {}
    for p in (emcgpu.FP64, emcgpu.FP32):
        ctx = _ctx(spec, n, p)
        ctx.init_particles_energy(0, 2025, 0.5, 0.05)
        ctx.update_screening(0)
        ctx.step(0.5e-15, ns, seed=2025)
        res[p] = (ctx.averages(0)["avgE"][:, 0], ctx.bath_history(0, 0), ctx.get_screening(0))
        ctx.close()
    e64, n64, q64 = res[emcgpu.FP64]
    e32, n32, q32 = res[emcgpu.FP32]
    print("screened cooling: <E> %.4f -> %.4f eV, <N> %.4f -> %.4f, q_s^2 %.3e (fp32 %.3e)" % (
        e64[0], e64[-1], n64[0], n64[-1], q64, q32), flush=True)
    assert e64[-1] < e64[0] - 0.03
    assert np.abs(e32 - e64).max() < 3e-3 * e64.max()
    assert np.abs(n32 - n64).max() < 1e-2 * n64.max()
    assert abs(q32 - q64) < 1e-2 * q64


def test_batched_screened_instances_are_independent():
    """Two plasma-screened hot-phonon instances with ragged counts in one
    arena (each with its own bath, screening and tables, the screening taken
    from its own tiles of the step partials) evolve exactly as each does
    alone."""
    emcgpu = _emc()
    import copy
    spec_a = MODELS["screened_hot_polar"]
    spec_b = copy.deepcopy(spec_a)
    spec_b["field_kvcm"] = 30.0
    spec_b["screening"]["volume"] = 1.0e-23
    counts = [3001, 1500]
    print("batched screened instances", counts, flush=True)
    ctx = emcgpu.Context(precision=emcgpu.FP64, n_instances=2)
    ctx.setup(0, spec_a)
    ctx.setup(1, spec_b)
    ctx.resize(counts)
    for i in range(2):
        ctx.init_particles_energy(i, 17, 0.5, 0.05)
        ctx.update_screening(i)
    ctx.step(0.5e-15, 40, seed=17)
    both = [(ctx.get_particles(i), ctx.bath_history(i, 0), ctx.get_screening(i)) for i in range(2)]
    ctx.close()
    for i, spec in enumerate((spec_a, spec_b)):
        # alone, at the same instance index (the Philox key includes it)
        solo = emcgpu.Context(precision=emcgpu.FP64, n_instances=2)
        solo.setup(0, spec_a)
        solo.setup(1, spec_b)
        solo.resize([counts[0] if i == 0 else 0, counts[1] if i == 1 else 0])
        solo.init_particles_energy(i, 17, 0.5, 0.05)
        solo.update_screening(i)
        solo.step(0.5e-15, 40, seed=17)
        p, h, q = solo.get_particles(i), solo.bath_history(i, 0), solo.get_screening(i)
        solo.close()
        assert (p["valley"] == both[i][0]["valley"]).all()
        assert np.array_equal(p["k"], both[i][0]["k"]) and np.array_equal(p["E"], both[i][0]["E"])
        assert np.array_equal(h, both[i][1]) and q == both[i][2]
    assert both[0][2] != both[1][2]
