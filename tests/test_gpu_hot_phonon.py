"""Hot-phonon path on the GPU: LDS-privatised bath tallies, the bath update
and the per-step table rebuild.

* Replay (FP64) against the reference's own hot-phonon run: the bath's
  mean occupation after every step and N_q every fifth step (golden
  fixtures recorded through emcPhononBath and the hot-phonon Froehlich
  mechanisms); the discrete history is covered with the other models by
  test_gpu_replay.
* Production (FP64) against the oracle's Philox twin: particle state, the
  bath after every step, the tallies of a record-only run (exact integers),
  reservoirs and lifetime profiles, and a rebuild stride.
* HIP-graph launch of hot-phonon steps equals eager launch; FP32 and FP64
  agree on the cooling curve.
"""
import numpy as np
import pytest

from emc_models import MODELS, field_force, photoexcited_cooling
from oracle import oracle_py as O

pytestmark = pytest.mark.gpu

REL_TOL = 1e-12


def _emc():
    from viennaemc_amd import emcgpu
    return emcgpu


def _rel_k(a, b):
    return np.linalg.norm(a - b, axis=-1) / np.maximum(np.linalg.norm(b, axis=-1), 1e-300)


def _ctx(spec, n, precision):
    emcgpu = _emc()
    ctx = emcgpu.Context(precision=precision)
    ctx.setup(0, spec)
    ctx.resize([n])
    return ctx


def test_replay_bath_matches_reference(golden):
    emcgpu = _emc()
    spec = MODELS["hot_polar"]
    b = "hot_polar.traj."
    np_, ns, seed = [int(x) for x in golden[b + "meta"]]
    dt = golden[b + "dt"][0]
    print("hot-phonon replay vs reference bath:", np_, "particles", ns, "steps", flush=True)
    draws = O.mt64_particle_draws(seed, np_, 2048)
    cursor = golden[b + "init_draws"].astype(np.uint64).copy()
    grain = golden[b + "init_grain"].copy()
    ctx = _ctx(spec, np_, emcgpu.FP64)
    ctx.set_particles(0, golden[b + "init_k"], golden[b + "init_E"], golden[b + "init_tau"],
                      golden[b + "init_valley"], golden[b + "init_sub"])
    mean, nq, events = [], [], []
    for s0 in range(0, ns, 5):
        tr = ctx.step(dt, 5, mode=emcgpu.REPLAY, step0=s0, draws=draws, cursor=cursor,
                      grain_tau=grain, trace=True)
        mean.append(ctx.bath_history(0, 0))
        nq.append(ctx.get_bath(0, 0)["nq"])
        events.append(tr["events"])
    mean = np.concatenate(mean)
    ref_mean = golden[b + "bath_mean"]
    ref_nq = golden[b + "bath_nq"].reshape(-1, spec["baths"][0]["n_bins"])
    assert (np.concatenate(events) == golden[b + "nevents"].reshape(ns, np_)).all()
    print("bath mean: max rel diff %.3e, exact %d / %d" % (
        (np.abs(mean - ref_mean) / ref_mean).max(), (mean == ref_mean).sum(), ns), flush=True)
    assert np.allclose(mean, ref_mean, rtol=1e-13, atol=0)
    assert np.allclose(np.array(nq), ref_nq, rtol=1e-13, atol=0)
    # the rebuilt device tables are the reference's tables for that bath
    M = O.OracleModel(spec)
    M.set_bath(0, nq[-1])
    M.L.orc_model_build_tables(M.h)
    cum, tau0 = ctx.get_tables(0, 0, spec["n_levels"])
    assert abs(tau0 - M.tau0(0)) <= 1e-13 * M.tau0(0)
    assert np.allclose(cum, M.cum(0), rtol=1e-13, atol=1e-16)


def _twin(spec, n, ns, seed, mode=(True, 1), sigma=0.05, step0=0):
    """GPU FP64 production run and the oracle twin, stepped together."""
    emcgpu = _emc()
    M = O.OracleModel(spec)
    M.set_bath_mode(update=mode[0], rebuild_every=mode[1], step_count=step0)
    ref = M.init_philox_energy(n, spec["init_energy"], sigma, seed)
    ctx = _ctx(spec, n, emcgpu.FP64)
    ctx.init_particles_energy(0, seed, spec["init_energy"], sigma)
    g0 = ctx.get_particles(0)
    assert _rel_k(g0["k"].reshape(-1, 3), ref["k"].reshape(-1, 3)).max() < REL_TOL
    assert (np.abs(g0["E"] - ref["E"]) / ref["E"]).max() < REL_TOL
    ctx.set_particles(0, ref["k"], ref["E"], ref["tau"], ref["valley"], ref["sub"])
    force, d = field_force(spec)
    ctx.step(0.5e-15, ns, seed=seed, step0=step0, bath_record_only=not mode[0],
             rebuild_every=mode[1])
    hist = []
    for s in range(ns):
        M.run_philox(ref, 1, 0.5e-15, force, d, seed, step0=step0 + s, obs=False)
        hist.append(M.bath(0)["mean"])
    return ctx, M, ref, np.array(hist)


def _check_state(ctx, ref):
    g = ctx.get_particles(0)
    assert (g["valley"] == ref["valley"]).all() and (g["sub"] == ref["sub"]).all()
    assert _rel_k(g["k"].reshape(-1, 3), ref["k"].reshape(-1, 3)).max() < REL_TOL
    assert (np.abs(g["E"] - ref["E"]) / ref["E"]).max() < REL_TOL


def test_production_bath_matches_oracle_twin():
    spec = MODELS["hot_polar"]
    print("hot-phonon production vs twin", flush=True)
    ctx, M, ref, hist = _twin(spec, 4096, 60, 0xB47B)
    _check_state(ctx, ref)
    gh = ctx.bath_history(0, 0)
    assert np.allclose(gh, hist, rtol=1e-12, atol=0)
    assert np.allclose(ctx.get_bath(0, 0)["nq"], M.bath(0)["nq"], rtol=1e-12, atol=0)
    assert hist[-1] != hist[0]  # the bath moved


def test_record_only_tallies_are_exact():
    spec = MODELS["hot_polar"]
    print("hot-phonon record-only tallies", flush=True)
    ctx, M, ref, _ = _twin(spec, 8192, 25, 0x7A11, mode=(False, 1))
    _check_state(ctx, ref)
    em, ab = ctx.bath_tallies(0, 0)
    ob = M.bath(0)
    assert em.sum() > 100 and ab.sum() > 10
    assert (em == ob["n_em"].astype(np.uint64)).all()
    assert (ab == ob["n_abs"].astype(np.uint64)).all()
    assert (ctx.get_bath(0, 0)["nq"] == ob["nq"]).all()  # no update ran


def test_rebuild_stride_matches_oracle_twin():
    spec = MODELS["hot_polar"]
    print("hot-phonon rebuild every 4th step", flush=True)
    ctx, M, ref, hist = _twin(spec, 4096, 22, 0x5712, mode=(True, 4), step0=3)
    _check_state(ctx, ref)
    assert np.allclose(ctx.bath_history(0, 0), hist, rtol=1e-12, atol=0)


def test_reservoirs_and_profile_match_oracle_twin():
    import copy
    spec = copy.deepcopy(MODELS["hot_polar"])
    b = spec["baths"][0]
    q = (np.arange(b["n_bins"]) + 0.5) * b["dq"]
    b.update(acoustic=True, ac_hw=0.03, tau_ac=5e-12, w_ridley=0.3, to_hw=0.02, tau_to=3e-12,
             tau_profile=1e-12 * (0.5 + q / (b["n_bins"] * b["dq"])), volume=3e-24)
    print("hot-phonon reservoirs + profile vs twin", flush=True)
    ctx, M, ref, hist = _twin(spec, 4096, 40, 0xAC0)
    _check_state(ctx, ref)
    g, o = ctx.get_bath(0, 0), M.bath(0)
    assert np.allclose(g["nq"], o["nq"], rtol=1e-12, atol=0)
    assert o["t_ac"] > 300.0 and o["t_to"] > 300.0  # both reservoirs heated
    for key in ("n_ac", "n_to", "t_ac", "t_to"):
        assert abs(g[key] - o[key]) <= 1e-12 * abs(o[key]), key
    assert np.allclose(ctx.bath_history(0, 0), hist, rtol=1e-12, atol=0)


@pytest.mark.parametrize("precision", [0, 1])
def test_graph_launch_with_baths_matches_eager(precision):
    emcgpu = _emc()
    spec = photoexcited_cooling(50_000)
    print("hot-phonon graph vs eager, precision", precision, flush=True)
    out = {}
    for graph in (False, True):
        ctx = _ctx(spec, 50_000, precision)
        ctx.init_particles_energy(0, 31, 0.5, 0.05)
        rec = []
        step0 = 0
        for ns in (30, 20, 45):
            ctx.step(0.5e-15, ns, seed=31, step0=step0, graph=graph)
            step0 += ns
            p = ctx.get_particles(0)
            rec.append((p["k"], p["E"], ctx.get_bath(0, 0)["nq"], ctx.bath_history(0, 0),
                        ctx.observables(0)))
        ctx.close()
        out[graph] = rec
    for a, b in zip(out[False], out[True]):
        for x, y in zip(a, b):
            assert np.array_equal(x, y)


def test_fp32_cooling_matches_fp64():
    """Photoexcited cooling (N(0.5, 0.05) eV, no field, 256-bin bath) for
    0.5 ps: FP32 and FP64 mean energy and mean phonon occupation agree."""
    emcgpu = _emc()
    n, ns = 400_000, 1000
    spec = photoexcited_cooling(n)
    res = {}
    for p in (emcgpu.FP64, emcgpu.FP32):
        ctx = _ctx(spec, n, p)
        ctx.init_particles_energy(0, 2024, 0.5, 0.05)
        ctx.step(0.5e-15, ns, seed=2024)
        res[p] = (ctx.averages(0)["avgE"][:, 0], ctx.bath_history(0, 0), ctx.get_bath(0, 0))
        ctx.close()
    e64, n64, b64 = res[emcgpu.FP64]
    e32, n32, b32 = res[emcgpu.FP32]
    print("cooling: <E> %.4f -> %.4f eV, <N> %.4f -> %.4f, max N_q %.3f" % (
        e64[0], e64[-1], n64[0], n64[-1], b64["nq"].max()), flush=True)
    assert e64[-1] < e64[0] - 0.05  # carriers cool by LO emission
    # ... into small-q phonons (q ~ k_i - k_f): those bins heat far above N0
    # while the q^2-weighted mean moves little
    assert b64["nq"].max() > 3 * b64["n0"]
    assert n64[-1] > n64[0]
    assert np.abs(e32 - e64).max() < 3e-3 * e64.max()
    assert np.abs(n32 - n64).max() < 1e-2 * n64.max()
    hot = b64["nq"] > b64["n0"] + 1.0
    assert hot.sum() >= 5
    assert np.allclose(b32["nq"][hot], b64["nq"][hot], rtol=0.1)


def test_bath_energy_record_and_balance():
    """The FP64 energy record of every bath update (emcPhononBath.hpp:264-358)
    against the oracle twin's, and the energy balance of photoexcited cooling
    (no field): per step the carriers' energy loss equals the net emitted
    LO phonon energy, which equals the LO energy gained plus the LO energy
    passed to the acoustic reservoir."""
    emcgpu = _emc()
    n, ns, seed = 4096, 50, 0xE4E
    spec = photoexcited_cooling(n)
    print("bath energy record vs twin", flush=True)
    M = O.OracleModel(spec)
    ref = M.init_philox_energy(n, spec["init_energy"], 0.05, seed)
    ctx = _ctx(spec, n, emcgpu.FP64)
    ctx.set_particles(0, ref["k"], ref["E"], ref["tau"], ref["valley"], ref["sub"])
    e_before = ref["E"].sum()
    ctx.step(0.5e-15, ns, seed=seed)
    force, d = field_force(spec)
    rec = []
    for s in range(ns):
        M.run_philox(ref, 1, 0.5e-15, force, d, seed, step0=s, obs=False)
        rec.append(M.bath_energy(0))
    rec = np.array(rec)
    got = ctx.bath_energy(0, 0)
    _check_state(ctx, ref)
    assert got.shape == (ns, 3)
    assert np.allclose(got, rec, rtol=1e-12, atol=1e-12 * np.abs(rec).max())
    # balance: net emission = LO gained + passed on (the clamp N_q >= 0 idle)
    assert np.allclose(got[:, 0], got[:, 1] + got[:, 2], rtol=0, atol=1e-9 * np.abs(got).max())
    # carriers: the energy lost in each step is the net emitted phonon energy
    sums = np.concatenate([[e_before], ctx.averages(0)["sumE"][:, 0]])
    loss = sums[:-1] - sums[1:]
    print("carrier loss %.6g eV, net emission %.6g eV, LO gained %.6g eV, passed on %.6g eV"
          % (loss.sum(), got[:, 0].sum(), got[:, 1].sum(), got[:, 2].sum()), flush=True)
    assert got[:, 0].sum() > 0
    assert np.allclose(loss, got[:, 0], rtol=0, atol=1e-10 * e_before)
