"""Replay-mode parity on the GPU (FP64): the HIP kernel consumes the
reference engine's recorded draws and must reproduce each particle's valley,
subvalley and scatter-event counts bit-exactly, and k and E within 1e-12
relative, against the reference's own trajectories
(golden fixtures) and, at larger sizes, against the oracle.
"""
import numpy as np
import pytest

from emc_models import MODELS, field_force
from oracle import oracle_py as O

pytestmark = pytest.mark.gpu

REL_TOL = 1e-12  # k and E, FP64 replay


def _emc():
    from viennaemc_amd import emcgpu
    return emcgpu


def _ctx(name, n):
    emcgpu = _emc()
    spec = MODELS[name]
    ctx = emcgpu.Context(precision=emcgpu.FP64)
    ctx.setup(0, spec)
    ctx.resize([n])
    return ctx


def _rel_k(a, b):
    na = np.linalg.norm(b, axis=-1)
    return np.linalg.norm(a - b, axis=-1) / np.maximum(na, 1e-300)


@pytest.mark.parametrize("name", list(MODELS))
def test_replay_matches_reference_trajectories(name, golden):
    emcgpu = _emc()
    spec = MODELS[name]
    b = name + ".traj."
    np_, ns, seed = [int(x) for x in golden[b + "meta"]]
    dt = golden[b + "dt"][0]
    print("replay vs reference:", name, np_, "particles", ns, "steps", flush=True)
    draws = O.mt64_particle_draws(seed, np_, 2048)
    cursor = golden[b + "init_draws"].astype(np.uint64).copy()
    grain = golden[b + "init_grain"].copy()
    ctx = _ctx(name, np_)
    ctx.set_particles(0, golden[b + "init_k"], golden[b + "init_E"], golden[b + "init_tau"],
                      golden[b + "init_valley"], golden[b + "init_sub"])
    tr = ctx.step(dt, ns, mode=emcgpu.REPLAY, draws=draws, cursor=cursor, grain_tau=grain,
                  trace=True, steps_per_launch=64)
    shape = (ns, np_)
    assert (tr["valley"] == golden[b + "valley"].reshape(shape)).all()
    assert (tr["sub"] == golden[b + "sub"].reshape(shape)).all()
    assert (tr["events"] == golden[b + "nevents"].reshape(shape)).all()
    total = golden[b + "init_draws"] + golden[b + "draws"].reshape(shape).sum(axis=0)
    assert (cursor == total).all()
    snaps = golden[b + "snap_steps"]
    kref = golden[b + "k"].reshape(len(snaps), np_, 3)
    Eref = golden[b + "E"].reshape(len(snaps), np_)
    assert _rel_k(tr["k"][snaps], kref).max() < REL_TOL
    assert (np.abs(tr["E"][snaps] - Eref) / np.abs(Eref)).max() < REL_TOL
    # the stored state after the last step agrees with the last trace row
    st = ctx.get_particles(0)
    assert (st["valley"] == tr["valley"][-1]).all()
    assert _rel_k(st["k"].reshape(-1, 3), kref[-1]).max() < REL_TOL


@pytest.mark.parametrize("name", ["si_jacoboni", "two_valley"])
def test_replay_matches_oracle_at_scale(name):
    """20k particles x 200 steps (config-2 physics) against the oracle."""
    emcgpu = _emc()
    spec = MODELS[name]
    n, ns, seed, dt = 20000, 200, 20261202, 1e-15
    print("replay vs oracle:", name, n, "particles", ns, "steps", flush=True)
    M = O.OracleModel(spec)
    draws = O.mt64_particle_draws(seed, n, 512)
    cur0 = np.zeros(n, dtype=np.uintp)
    st = M.init_replay(draws, cur0, O.thermal_voltage(spec["T"]))
    ctx = _ctx(name, n)
    ctx.set_particles(0, st["k"], st["E"], st["tau"], st["valley"], st["sub"])
    cursor = cur0.astype(np.uint64).copy()
    grain = st["grain"].copy()
    force, _ = field_force(spec)
    ctx.step(dt, ns, mode=emcgpu.REPLAY, draws=draws, cursor=cursor, grain_tau=grain,
             steps_per_launch=50)
    g = ctx.get_particles(0)
    ocur = cur0.copy()
    nover, _ = M.run_replay(st, draws, ocur, ns, dt, force)
    assert nover == 0
    assert (g["valley"] == st["valley"]).all()
    assert (g["sub"] == st["sub"]).all()
    assert (cursor == ocur.astype(np.uint64)).all()
    assert _rel_k(g["k"].reshape(-1, 3), st["k"].reshape(-1, 3)).max() < REL_TOL
    assert (np.abs(g["E"] - st["E"]) / np.abs(st["E"])).max() < REL_TOL


def test_replay_reports_exhausted_draws():
    emcgpu = _emc()
    name = "si_jacoboni"
    spec = MODELS[name]
    n = 512
    M = O.OracleModel(spec)
    draws = O.mt64_particle_draws(7, n, 16)
    cur = np.zeros(n, dtype=np.uintp)
    st = M.init_replay(draws, cur, O.thermal_voltage(spec["T"]))
    ctx = _ctx(name, n)
    ctx.set_particles(0, st["k"], st["E"], st["tau"], st["valley"], st["sub"])
    cursor = cur.astype(np.uint64).copy()
    with pytest.raises(emcgpu.EmcGpuError) as e:
        ctx.step(1e-15, 200, mode=emcgpu.REPLAY, draws=draws, cursor=cursor)
    assert e.value.status == 4
