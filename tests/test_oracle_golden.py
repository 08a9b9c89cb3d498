"""Pin the oracle (oracle/emc_oracle.c) to the reference's own outputs.

The golden vectors in tests/golden/reference_golden.npz were recorded by
oracle/ref_harness.cpp, which drives the reference's classes compiled from
the ViennaEMC headers where they lie (see tests/golden/make_golden.py).
Everything here is bit-exact: the oracle restates the reference with the
same FP64 operation order, built by gcc without FMA contraction.
"""
import numpy as np
import pytest

from emc_models import MODELS, field_force
from oracle import oracle_py as O

NAMES = list(MODELS)


def _sm64(x):
    m = (1 << 64) - 1
    z = (x + 0x9E3779B97F4A7C15) & m
    z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & m
    z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & m
    return z ^ (z >> 31)


# ---- random numbers ----------------------------------------------------

def test_mt19937_64_known_answer(golden):
    # [rand.predef]: the 10000th output of a default-constructed mt19937_64
    assert O.mt64_fill(5489, 10000)[-1] == 9981545732273789042
    assert golden["mt64.kat10000"][0] == 9981545732273789042
    assert (O.mt64_fill(20261202, 8) == golden["mt64.seed20261202"]).all()


def test_uniform_real_distribution_mapping(golden):
    raw = O.mt64_fill(99, 8)
    ref = golden["mt64.uniform99"]
    for i in range(8):
        c = O.canonical(int(raw[i]))
        u = c * (1.0 - 1e-6) + 1e-6 if i % 2 else c
        assert u == ref[i]


def test_canonical_clamps_below_one():
    # 2^64 - 1 rounds to 2^64 as a double: generate_canonical clamps to nextafter(1, 0)
    assert O.canonical((1 << 64) - 1) == np.nextafter(1.0, 0.0)
    assert O.canonical(0) == 0.0
    assert O.canonical(1 << 63) == 0.5


def test_splitmix_particle_seed():
    # the harness seeds particle p with splitmix64(seed ^ p)
    for seed, pid in [(20261202, 0), (20261202, 17), (1, 1 << 40)]:
        assert O.particle_seed(seed, pid) == _sm64(seed ^ pid)


def test_philox_known_answers():
    # Random123 known-answer vectors for philox4x32-10
    assert list(O.philox([0, 0, 0, 0], [0, 0])) == [0x6627E8D5, 0xE169C58D, 0xBC57AC4C, 0x9B00DBD8]
    f = 0xFFFFFFFF
    assert list(O.philox([f, f, f, f], [f, f])) == [0x408F276D, 0x41C83B0E, 0xA20BC7C6, 0x6D5451FD]
    assert list(O.philox([0x243F6A88, 0x85A308D3, 0x13198A2E, 0x03707344],
                         [0xA4093822, 0x299F31D0])) == [0xD16CFE09, 0x94FDCCEB, 0x5001E420, 0x24126EA1]


def test_production_uniform_layout():
    """Uniform j of a production stream is word j % 4 of Philox block
    (j / 4, step, particle, instance) under the seed's two halves, its top 24
    bits forced odd: strictly inside (0, 1) and exact in FP32, which is what
    lets the FP32 and FP64 kernels draw the same numbers."""
    import numpy as np
    seed = 0x0123456789ABCDEF
    key = [seed & 0xFFFFFFFF, seed >> 32]
    for j, step, pid, inst in [(0, 0, 0, 0), (3, 7, 11, 2), (4, 0xFFFFFFFF, 5, 0),
                               (9, 123456, 987654, 0x40000001)]:
        w = O.philox([j >> 2, step, pid, inst], key)[j & 3]
        u = O.prod_uniform(seed, inst, pid, step, j)
        assert u == ((int(w) >> 8) | 1) / 2.0 ** 24
        assert 0.0 < u < 1.0 and float(np.float32(u)) == u
    # the extremes of the 24-bit grid
    assert O.load().orc_philox_u01_24(0) == 2.0 ** -24
    assert O.load().orc_philox_u01_24(0xFFFFFFFF) == 1.0 - 2.0 ** -24


# ---- tables --------------------------------------------------------------

@pytest.mark.parametrize("name", NAMES)
def test_rates_and_tables_bit_exact(name, golden, oracle_models):
    spec = MODELS[name]
    M = oracle_models[name]
    nm = len(spec["mechs"])
    en = golden[name + ".rate_energies"]
    rs = golden[name + ".rate_sample"].reshape(nm, -1)
    rt = golden[name + ".rate_table"].reshape(nm, -1)
    dE = spec["max_energy"] / spec["n_levels"]
    for j in range(nm):
        assert all(M.rate(j, e) == r for e, r in zip(en, rs[j])), (name, j)
        assert all(M.rate(j, (l + 1) * dE) == rt[j, l] for l in range(spec["n_levels"])), (name, j)
    for v in range(M.n_valleys):
        assert M.tau0(v) == golden[name + ".tau0"][v]


@pytest.mark.parametrize("name", NAMES)
def test_tables_invariants(name, oracle_models):
    """Cumulative tables are non-decreasing over mechanisms, max of the last is 1."""
    M = oracle_models[name]
    for v in range(M.n_valleys):
        cum = M.cum(v)
        assert (np.diff(cum, axis=0) >= 0).all()
        assert cum[-1].max() == 1.0
        assert (cum >= 0).all()


def test_energy_level_quirk(oracle_models):
    """getEnergyLevel: floor(E/dE) - 1, -1 maps to 0, clamps high (emcScatterHandler.hpp:237-244)."""
    M = oracle_models["si_jacoboni"]
    dE = 2.0 / 2000
    assert M.energy_level(0.0) == 0
    assert M.energy_level(0.5 * dE) == 0
    assert M.energy_level(1.0 * dE) == 0
    assert M.energy_level(2.0 * dE) == 1
    assert M.energy_level(2.5 * dE) == 1
    assert M.energy_level(5.0) == 1999
    # a negative energy wraps size_t and lands on the last level
    assert M.energy_level(-3 * dE) == 1999


# ---- single operations ------------------------------------------------------

@pytest.mark.parametrize("name", NAMES)
def test_drift_velocity_bit_exact(name, golden, oracle_models):
    M = oracle_models[name]
    rows = golden[name + ".drift"].reshape(-1, 20)
    for r in rows:
        v, s = int(r[0]), int(r[1])
        k, E, pos = M.drift(v, s, r[8], r[2:5], 0.0, [1e-9, 2e-9, 3e-9], r[5:8])
        assert (k == r[9:12]).all() and E == r[12] and (pos == r[13:16]).all()
        assert (M.velocity(v, s, k, E) == r[16:19]).all()
        assert M.normk(v, E) == r[19]


@pytest.mark.parametrize("name", NAMES)
def test_final_states_bit_exact(name, golden, oracle_models):
    """Every mechanism's scatterParticle: draws consumed, valley, subvalley, k,
    E; hot-phonon mechanisms also the bath bin they record |k' - k| in."""
    M = O.OracleModel(MODELS[name]) if MODELS[name].get("baths") else oracle_models[name]
    rows = golden[name + ".final"].reshape(-1, 14)
    bins = golden[name + ".final_bins"].reshape(-1, 2) if M.n_baths else None
    for idx, r in enumerate(rows):
        j, t = int(r[0]), idx % 48
        draws = O.mt64_fill(_sm64(1000 + t + 7919 * j), 16)
        b0 = M.bath(0) if M.n_baths else None
        k, E, vv, ss, used = M.mech_scatter(j, r[3:6], r[6], int(r[1]), int(r[2]), draws)
        assert used == int(r[7]), (name, idx)
        assert (vv, ss) == (int(r[8]), int(r[9])), (name, idx)
        assert (k == r[10:13]).all() and E == r[13], (name, idx)
        if bins is not None:
            b1 = M.bath(0)
            dem = np.nonzero(b1["n_em"] != b0["n_em"])[0]
            dab = np.nonzero(b1["n_abs"] != b0["n_abs"])[0]
            if bins[idx, 1] == 1:
                assert list(dem) == [bins[idx, 0]] and len(dab) == 0, (name, idx)
            else:
                assert list(dab) == [bins[idx, 0]] and len(dem) == 0, (name, idx)


# ---- trajectories -----------------------------------------------------------

@pytest.mark.parametrize("name", NAMES)
def test_replay_trajectories_bit_exact(name, golden, oracle_models):
    """Per-particle reference streams: init, 300 steps, every event and state.
    Hot-phonon models also update the bath and rebuild the tables after
    every step; the bath's mean occupation is compared step by step and its
    N_q every fifth step."""
    spec = MODELS[name]
    M = O.OracleModel(spec)  # fresh: a bath evolves with the run
    b = name + ".traj."
    np_, ns, seed = [int(x) for x in golden[b + "meta"]]
    dt = golden[b + "dt"][0]
    force = golden[b + "force"]
    assert (np.array(field_force(spec)[0]) == force).all()
    draws = O.mt64_particle_draws(seed, np_, 4096)
    cur = np.zeros(np_, dtype=np.uintp)
    st = M.init_replay(draws, cur, O.thermal_voltage(spec["T"]))
    for key, g in (("k", "init_k"), ("E", "init_E"), ("tau", "init_tau"), ("grain", "init_grain"),
                   ("valley", "init_valley"), ("sub", "init_sub")):
        assert (st[key] == golden[b + g]).all(), key
    assert (cur == golden[b + "init_draws"]).all()
    # positions in the periodic device box, started on the harness's fixed set
    M.set_box(golden[b + "box"])
    st["pos"] = golden[b + "init_pos"].copy()
    plasma = spec.get("screening", {}).get("plasma", False)
    if plasma:
        # hotCarrierMHP seeds the screening from the initial carriers
        M.seed_screening(st["E"], st["valley"])
    if M.n_baths:
        tr = {key: [] for key in ("valley", "sub", "nev", "k", "E", "pos")}
        mean, nq, qs2 = [], [], []
        for s in range(ns):
            nover, t1 = M.run_replay(st, draws, cur, 1, dt, force, trace=True)
            assert nover == 0
            for key in tr:
                tr[key].append(t1[key])
            bt = M.bath(0)
            mean.append(bt["mean"])
            qs2.append(M.qs2())
            if s % 5 == 4:
                nq.append(bt["nq"])
        tr = {key: np.concatenate(v) for key, v in tr.items()}
        assert (np.array(mean) == golden[b + "bath_mean"]).all()
        assert (np.concatenate(nq) == golden[b + "bath_nq"]).all()
        if plasma:
            # q_s^2 after every step's updateScreening
            assert (np.array(qs2) == golden[b + "qs2"]).all()
    else:
        nover, tr = M.run_replay(st, draws, cur, ns, dt, force, trace=True)
        assert nover == 0
    shape = (ns, np_)
    assert (tr["valley"].reshape(shape) == golden[b + "valley"].reshape(shape)).all()
    assert (tr["sub"].reshape(shape) == golden[b + "sub"].reshape(shape)).all()
    assert (tr["nev"].reshape(shape) == golden[b + "nevents"].reshape(shape)).all()
    snaps = golden[b + "snap_steps"]
    assert (tr["k"].reshape(ns, np_, 3)[snaps].reshape(-1) == golden[b + "k"]).all()
    assert (tr["E"].reshape(ns, np_)[snaps].reshape(-1) == golden[b + "E"]).all()
    # driftParticle of the bulk handler: position += dt * velocity, wrapped
    pos = tr["pos"].reshape(ns, np_, 3)
    assert (pos[snaps].reshape(-1) == golden[b + "pos"]).all()
    box = golden[b + "box"]
    assert (pos >= 0).all() and (pos <= box).all()
    total = golden[b + "init_draws"] + golden[b + "draws"].reshape(shape).sum(axis=0)
    assert (cur == total).all()


@pytest.mark.parametrize("name", NAMES)
def test_bulk_handler_loop_bit_exact(name, golden, oracle_models):
    """The real basicBulkParticleHandler (shared stream): particle creation,
    moveParticles and the per-step observables of bulkSimulation.cpp."""
    spec = MODELS[name]
    M = O.OracleModel(spec)  # fresh: a bath evolves with the run
    hb = name + ".handler."
    n, nsteps, seed, nodes = [int(x) for x in golden[hb + "meta"]]
    dt, doping, spacing, T = golden[hb + "geom"]
    mp = spacing * (nodes - 1)
    n2, obs, hist = M.run_shared(seed, [mp] * 3, [spacing] * 3, doping, T, spec["field_dir"],
                                 spec["field_kvcm"] * 1e5, dt, nsteps, bath_history=True)
    assert n2 == n
    if M.n_baths:
        # hotPhononGa2O3's loop: moveParticles, bath update, table rebuild
        assert (hist[:, 0] == golden[hb + "bath_mean"]).all()
    ref = golden[hb + "obs"].reshape(nsteps + 1, 3, M.n_valleys)
    assert (obs == ref).all()


def _bulk_dump(box, st, pos):
    """basicBulkParticleHandler::print (basicBulkParticleHandler.hpp:227-248)
    in the reference's default stream format (%g)"""
    lines = [" ".join("%g" % x for x in box)]
    k = st["k"].reshape(-1, 3)
    pos = pos.reshape(-1, 3)
    for i in range(st["E"].size):
        lines.append("%d %s %s %g %d %d" % (i, " ".join("%g" % x for x in pos[i]),
                                            " ".join("%g" % x for x in k[i]), st["E"][i],
                                            st["sub"][i], st["valley"][i]))
    return "\n".join(lines).encode()


@pytest.mark.parametrize("name", [n for n in NAMES if not MODELS[n].get("baths")])
def test_bulk_handler_files(name, golden):
    """The bulk example's files of the real handler (bulkSimulation.cpp:
    127-128, 171-183): the oracle's particles with their positions in the
    periodic box after generateInitialParticles and after the last step, and
    the three time series, printed in the reference's format, equal the
    reference's files byte for byte."""
    spec = MODELS[name]
    hb = name + ".handler."
    n, nsteps, seed, nodes = [int(x) for x in golden[hb + "meta"]]
    dt, doping, spacing, T = golden[hb + "geom"]
    box = [spacing * (nodes - 1)] * 3
    args = ([spacing] * 3, doping, T, spec["field_dir"], spec["field_kvcm"] * 1e5, dt)
    _, _, eq = O.OracleModel(spec).run_shared(seed, box, *args, 0, state=True)
    _, obs, end = O.OracleModel(spec).run_shared(seed, box, *args, nsteps, state=True)
    assert (eq["pos"] == end["pos0"]).all() and (eq["pos0"] == eq["pos"]).all()
    pos = end["pos"].reshape(-1, 3)
    assert (pos >= 0).all() and (pos <= box[0]).all()
    assert _bulk_dump(box, eq, eq["pos"]) == bytes(golden["bulkfile.%s.ElectronsEq" % name])
    assert _bulk_dump(box, end, end["pos"]) == bytes(golden["bulkfile.%s.ElectronsEnd" % name])
    for j, part in enumerate(("AvgEnergy", "AvgDriftVelocity", "valleyOccupation")):
        text = "".join("%g %s\n" % (s * dt, " ".join("%g" % x for x in obs[s, j]))
                       for s in range(nsteps + 1))
        assert text.encode() == bytes(golden["bulkfile.%s.%s" % (name, part)]), part
