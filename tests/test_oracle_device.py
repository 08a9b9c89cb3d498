"""The oracle's 2-D device path against the reference's own device classes.

Golden fixtures recorded by the harness through emcDevice, the NGP / CIC /
NEC particle-mesh schemes, calcEField, emcSORSolver and
emcBasicParticleHandler (a 21 x 7 node n+ n n+ resistor, ohmic contacts at
both ends):
* charge assignment, force interpolation and E-field: bit-exact;
* SOR equilibrium and non-equilibrium solutions: bit-exact;
* 64 particles x 150 steps of the device particle loop under a frozen field
  with per-particle streams (drift with positions, boundary reflection,
  removal at contacts): bit-exact;
* full simulations in the emcSimulation step order with one shared stream
  (equilibrium, initial particles from the doping or, with
  usePotentialForInit, from the equilibrium potential, Poisson every step
  or every third step,
  drift/scatter, ohmic contact equilibrium, charge assignment, the
  drift-velocity channel current), CIC and NGP: bit-exact;
* the reference's own emcSimulation::execute (Poisson every second step,
  transient steps, final averages): its averaged channel current, contact
  current file and averaged potential file from the oracle's run.
* the 3-D NGP and CIC schemes and the grid-point E-field (an 8 x 6 x 5 grid
  with a different spacing per dimension, ohmic contacts over parts of two
  faces): bit-exact.
The mathematics pins what the fixtures leave open: CIC weights of every
particle sum to one, NGP mass lands on the nearest node, the SOR fixed point
solves the discrete Poisson equation.
"""
import numpy as np
import pytest

from emc_models import DEVICE_GATED, DEVICE_SMALL, MODELS_DEVICE, device_region_specs
from oracle import oracle_py as O


@pytest.fixture(scope="module")
def dev():
    return O.OracleDevice(DEVICE_SMALL)


def _model():
    return O.OracleModel(MODELS_DEVICE["device_para"])


def test_device_geometry(dev):
    assert dev.ext == (21, 7)
    g = dev.grids()
    d = g["doping"].reshape(7, 21)
    assert (d[:, :4] == 1e25).all() and (d[:, 17:] == 1e25).all() and (d[:, 4:17] == 1e24).all()
    c = g["contact"].reshape(7, 21)
    assert (c[:, 0] == 0).all() and (c[:, -1] == 1).all() and (c[1:-1, 1:-1] == -1).all()
    assert abs(dev.Vt - 1.38066e-23 / 1.60219e-19 * 300) < 1e-18
    assert abs(dev.cell_volume - 1e-9 * 1e-9 * 1e-6) < 1e-40


@pytest.mark.parametrize("scheme,name", [(O.PM_NGP, "ngp"), (O.PM_CIC, "cic"), (O.PM_NEC, "nec")])
def test_pm_schemes_bit_exact(dev, golden, scheme, name):
    pos = golden["device.kat.pos"]
    g = dev.assign(scheme, pos)
    assert (g == golden["device.kat.assign_" + name]).all()
    f = dev.forces(scheme, golden["device.kat.ex"], golden["device.kat.ey"], pos)
    assert (f.reshape(-1) == golden["device.kat.force_" + name]).all()
    # every particle deposits exactly its carrier count
    assert abs(g.sum() - pos.size // 2) < 1e-12


def test_cic_weights_are_the_reference_ones(dev):
    """The reference gives node (floor x, floor y) the weight wX * wY (the
    mirror of the textbook CIC weight); a particle one quarter into a cell
    puts 1/16 on the lower node and 9/16 on the upper one."""
    pos = np.array([[0.25e-9 + 5e-9, 0.25e-9 + 2e-9]])
    g = dev.assign(O.PM_CIC, pos).reshape(7, 21)
    assert abs(g[2, 5] - 0.0625) < 1e-15 and abs(g[3, 6] - 0.5625) < 1e-15
    assert abs(g[2, 6] - 0.1875) < 1e-15 and abs(g[3, 5] - 0.1875) < 1e-15
    ngp = dev.assign(O.PM_NGP, np.array([[5.4e-9, 2.6e-9]])).reshape(7, 21)
    assert ngp[3, 5] == 1.0 and ngp.sum() == 1.0


def test_efield_bit_exact(dev, golden):
    pot = golden["device.kat.pot"]
    ex, ey = dev.efield(O.PM_CIC, pot)
    assert (ex == golden["device.kat.efield_grid_x"]).all()
    assert (ey == golden["device.kat.efield_grid_y"]).all()
    ex, ey = dev.efield(O.PM_NEC, pot)
    assert (ex == golden["device.kat.efield_edge_x"]).all()
    assert (ey == golden["device.kat.efield_edge_y"]).all()
    # central difference in the interior
    ex, _ = dev.efield(O.PM_CIC, pot)
    p = pot.reshape(7, 21)
    want = (p[3, 4] - p[3, 6]) * dev.Vt / (2 * 1e-9)
    assert abs(ex.reshape(7, 21)[3, 5] - want) <= 1e-15 * abs(want)


@pytest.mark.parametrize("scheme,name", [(O.PM_NGP, "ngp"), (O.PM_CIC, "cic")])
def test_pm3_schemes_bit_exact(golden, scheme, name):
    """3-D assignToMesh and interpolateForce (emcCICScheme.hpp:52-74,
    143-173; emcNGPScheme.hpp:24-66) against the reference's values."""
    ext, h = golden["pm3.ext"].astype(np.int32), golden["pm3.spacing"]
    pos = golden["pm3.pos"].reshape(-1, 3)
    g = O.pm3_assign(scheme, ext, h, pos)
    assert (g == golden["pm3.assign_" + name]).all()
    assert abs(g.sum() - pos.shape[0]) < 1e-12  # one carrier per particle
    f = O.pm3_force(scheme, ext, h, golden["pm3.ex"], golden["pm3.ey"], golden["pm3.ez"], pos,
                    -O.Q)
    assert (f.reshape(-1) == golden["pm3.force_" + name]).all()


def test_pm3_weights_and_efield():
    """What the mathematics fixes: a CIC particle a quarter into a cell
    along every axis puts (1/4)^3 on the node at the floor corner (the
    reference's mirrored weight) and (3/4)^3 on the opposite one; NGP puts
    its carrier on the nearest node; a linear potential gives a constant
    interior field -dphi/dx * vt, zero at artificial faces and the inner
    value at contacts."""
    ext, h = np.array([5, 4, 3], dtype=np.int32), np.array([1e-9, 2e-9, 0.5e-9])
    g = O.pm3_assign(O.PM_CIC, ext, h, np.array([[1.25e-9, 2.5e-9, 0.125e-9]])).reshape(3, 4, 5)
    assert abs(g[0, 1, 1] - 0.25 ** 3) < 1e-16 and abs(g[1, 2, 2] - 0.75 ** 3) < 1e-16
    assert abs(g.sum() - 1.0) < 1e-15
    n = O.pm3_assign(O.PM_NGP, ext, h, np.array([[2.4e-9, 3.1e-9, 0.7e-9]])).reshape(3, 4, 5)
    assert n[1, 2, 2] == 1.0 and n.sum() == 1.0
    nn = int(np.prod(ext))
    kk, jj, ii = np.meshgrid(range(3), range(4), range(5), indexing="ij")
    pot = (0.3 * ii).reshape(-1).astype(np.float64)  # slope 0.3 per node along x
    bc = np.full((3, nn), -2, dtype=np.int32)
    for d, c in ((0, ii), (1, jj), (2, kk)):
        on = ((c == 0) | (c == ext[d] - 1)).reshape(-1)
        bc[d, on] = -1
    xmin = (ii == 0).reshape(-1)
    bc[0, xmin] = 0  # XMIN a contact
    ex, ey, ez = O.pm3_efield(ext, h, 0.025, bc, pot)
    want = -0.3 * 0.025 / 1e-9
    inner = ((ii > 0) & (ii < 4)).reshape(-1)
    assert np.allclose(ex[inner], want, rtol=1e-13)
    assert np.allclose(ex[xmin], want, rtol=1e-13)      # contact: inner neighbour
    assert (ex[(ii == 4).reshape(-1)] == 0).all()       # artificial XMAX
    assert (ey == 0).all() and (ez == 0).all()


def test_pm3_efield_bit_exact(golden):
    """3-D calcEFieldAtGridPts with setEFieldBoundaryValues (emcEFieldCalculation.hpp:
    13-32, 57-81) against the reference's field, the faces' contact status
    as the reference's emcSurface reports it."""
    ext, h = golden["pm3.ext"].astype(np.int32), golden["pm3.spacing"]
    E = O.pm3_efield(ext, h, float(golden["pm3.vt"][0]), golden["pm3.bc"], golden["pm3.pot"])
    for d, c in enumerate("xyz"):
        assert (E[d] == golden["pm3.efield_" + c]).all()


def test_sor_bit_exact(dev, golden):
    pot0 = dev.init_potential()
    eq, sweeps = dev.sor(pot0, accuracy=1e-12, omega=1.8)
    assert sweeps > 10
    assert (eq == golden["device.kat.sor_eq"]).all()
    neq, _ = dev.sor(eq, conc=golden["device.kat.sor_conc"], accuracy=1e-12, omega=1.8)
    assert (neq == golden["device.kat.sor_neq"]).all()


def test_sor_fixed_point_solves_poisson(dev, golden):
    """At the converged point the discrete equation
    sum_d hF_d (phi_prev + phi_next - 2 phi) + hprod (exp(-phi) - n + C) = 0
    holds at every non-contact node (Neumann by mirrored neighbours)."""
    phi = golden["device.kat.sor_neq"].reshape(7, 21)
    n = golden["device.kat.sor_conc"].reshape(7, 21)
    C = dev.grids()["doping"].reshape(7, 21) / DEVICE_SMALL["ni"]
    h = 1e-9 / dev.debye
    worst = 0.0
    for iy in range(7):
        for ix in range(1, 20):
            xm = phi[iy, ix - 1]
            xp = phi[iy, ix + 1]
            ym = phi[iy - 1, ix] if iy > 0 else phi[iy + 1, ix]
            yp = phi[iy + 1, ix] if iy < 6 else phi[iy - 1, ix]
            r = (xm + xp + ym + yp - 4 * phi[iy, ix]) + h * h * (np.exp(-phi[iy, ix]) - n[iy, ix] + C[iy, ix])
            worst = max(worst, abs(r) / (h * h * abs(C[iy, ix])))
    assert worst < 1e-9


def test_sor_with_gates_bit_exact(golden):
    """Two gate contacts (YMAX over x = 5..15 nm, YMIN over 8..12 nm) add the
    Robin terms of checkForGateContact: barrier height only in the
    equilibrium solve, barrier plus gate voltage in the non-equilibrium one."""
    cfg = golden["device.gate.cfg"]
    gates = [c for c in DEVICE_GATED["contacts"] if len(c) == 7]
    assert np.array_equal(cfg, np.array([x for g in gates for x in (g[1],) + tuple(g[2:])]))
    dev = O.OracleDevice(DEVICE_GATED)
    c = dev.grids()["contact"].reshape(7, 21)
    # a gate is no Dirichlet contact
    assert (c[:, 0] == 0).all() and (c[:, -1] == 1).all() and (c[:, 1:-1] == -1).all()
    eq, _ = dev.sor(dev.init_potential(), accuracy=1e-12, omega=1.8)
    assert (eq == golden["device.gate.sor_eq"]).all()
    neq, _ = dev.sor(eq, conc=golden["device.gate.sor_conc"], accuracy=1e-12, omega=1.8)
    assert (neq == golden["device.gate.sor_neq"]).all()
    # without the gates the solution is another one
    assert np.abs(eq - golden["device.kat.sor_eq"]).max() > 1.0


def test_sor_gate_fixed_point_is_the_robin_equation(golden):
    """At a gate node of boundary YMAX the converged non-equilibrium point
    solves the Poisson equation whose missing y-neighbour is the mirrored one
    plus the oxide flux: hF_y (2 phi_prev - 2 phi) + g h hF_y (V_g - phi),
    g = 2 (eps_ox / eps_sc) / (t_ox / L_D), V_g = (barrier + voltage) / V_t;
    a node of the boundary outside the gate keeps the plain Neumann form."""
    dev = O.OracleDevice(DEVICE_GATED)
    phi = golden["device.gate.sor_neq"].reshape(7, 21)
    n = golden["device.gate.sor_conc"].reshape(7, 21)
    C = dev.grids()["doping"].reshape(7, 21) / DEVICE_GATED["ni"]
    h = 1e-9 / dev.debye
    worst = {}
    for g in [c for c in DEVICE_GATED["contacts"] if len(c) == 7]:
        b, v, lo, hi, eps_ox, t_ox, barrier = g
        iy, iy_in = (6, 5) if b == 3 else (0, 1)
        gf = 2 * (eps_ox / DEVICE_GATED["eps_r"]) / (t_ox / dev.debye)
        vg = (barrier + v) / dev.Vt
        for ix in range(1, 20):
            on_gate = round(lo / 1e-9) <= ix <= round(hi / 1e-9)
            r = (phi[iy, ix - 1] + phi[iy, ix + 1] - 2 * phi[iy, ix]
                 + 2 * phi[iy_in, ix] - 2 * phi[iy, ix]
                 + h * h * (np.exp(-phi[iy, ix]) - n[iy, ix] + C[iy, ix]))
            diag = 4 + h * h * (np.exp(-phi[iy, ix]) + n[iy, ix])
            if on_gate:
                r += gf * h * (vg - phi[iy, ix])
                diag += gf * h
            # the residual as a potential change: within 100 x the 1e-12
            # stopping rule (a dropped or mis-signed gate term is ~1)
            worst[on_gate] = max(worst.get(on_gate, 0.0), abs(r) / diag)
    assert worst[True] < 1e-10 and worst[False] < 1e-10


def test_device_trajectories_bit_exact(dev, golden):
    M = _model()
    b = "device.traj."
    np_, ns, seed = [int(x) for x in golden[b + "meta"]]
    dt = golden[b + "dt"][0]
    draws = O.mt64_particle_draws(seed, np_, 512)
    cursor = np.zeros(np_, dtype=np.uintp)
    coords = [((p * 5) % 21, p % 7) for p in range(np_)]
    st = dev.init_replay(M, coords, draws, cursor)
    init = golden[b + "init"].reshape(np_, 8)
    assert (st[:, :8] == init).all()
    assert (cursor == golden[b + "init_draws"]).all()
    rem, con, tr = dev.run_replay(M, O.PM_CIC, golden[b + "ex"], golden[b + "ey"], st, draws, cursor,
                                  ns, dt)
    assert (rem == golden[b + "removed_step"]).all()
    who = golden[b + "removed_who"]
    assert (con[who] == golden[b + "removed_contact"]).all()
    assert (tr["nev"] == golden[b + "nevents"]).all()
    assert (tr["pos"] == golden[b + "pos"]).all()
    snaps = golden[b + "snap_steps"]  # k and E are kept every fifth step
    assert (tr["k"].reshape(ns, np_, 3)[snaps].reshape(-1) == golden[b + "k"]).all()
    assert (tr["E"].reshape(ns, np_)[snaps].reshape(-1) == golden[b + "E"]).all()
    total = golden[b + "init_draws"] + golden[b + "draws"].reshape(ns, np_).sum(axis=0)
    assert (cursor == total).all()


# drift-velocity channel current region of the harness runs [m]
CHANNEL = (5e-9, 15e-9, 10e-9)


@pytest.mark.parametrize("pm,tag", [(O.PM_CIC, "cic"), (O.PM_NGP, "ngp"), (O.PM_CIC, "cic_sub3"),
                                    (O.PM_CIC, "cic_potinit")])
def test_device_simulation_bit_exact(dev, golden, pm, tag):
    b = "device." + tag + "."
    seed, nsteps, dt, acc, omega, interval, pot_init = golden[b + "cfg"]
    out = dev.run(_model(), pm, acc, omega, dt, int(nsteps), int(seed), interval=int(interval),
                  channel=CHANNEL, init_from_potential=bool(pot_init))
    # getChannelDriftCurrent after every step
    assert (out["chan"] == golden[b + "chan"]).all()
    if interval > 1:  # the potential changes only on solve steps
        pot = out["pot"]
        for s in range(1, int(nsteps)):
            assert (s % int(interval) == 0) != (pot[s] == pot[s - 1]).all(), s
    assert (out["eq_pot"] == golden[b + "eqpot"]).all()
    assert out["n_init"][0] == golden[b + "ninit"][0]
    assert (out["pot"].reshape(-1) == golden[b + "pot"]).all()
    assert (out["n_part"] == golden[b + "npart"]).all()
    assert (out["exits"].reshape(-1) == golden[b + "exits"]).all()
    # handleOhmicContacts returns injected - deleted per contact
    assert ((out["injs"] - out["dels"]).reshape(-1) == golden[b + "ohmic"]).all()
    assert (out["nr_part"] == golden[b + "nrpart_last"]).all()
    assert out["exits"].sum() > 20 and out["injs"].sum() > 20


def test_device_simulation_with_region_tables_bit_exact(golden):
    """Mechanisms registered for some doping regions only: each region has its
    own (valley, region) tables and tau0, and a particle scatters with the
    tables of the region of its last drift."""
    b = "device.cic_regions."
    seed, nsteps, dt, acc, omega, interval, pot_init = golden[b + "cfg"]
    dev = O.OracleDevice(DEVICE_SMALL)
    models = [O.OracleModel(sp) for sp in device_region_specs()]
    assert len({M.tau0(0) for M in models}) == 3
    dev.set_region_models(models)
    out = dev.run(models[0], O.PM_CIC, acc, omega, dt, int(nsteps), int(seed), channel=CHANNEL)
    assert (out["chan"] == golden[b + "chan"]).all()
    assert out["n_init"][0] == golden[b + "ninit"][0]
    assert (out["pot"].reshape(-1) == golden[b + "pot"]).all()
    assert (out["n_part"] == golden[b + "npart"]).all()
    assert (out["exits"].reshape(-1) == golden[b + "exits"]).all()
    assert ((out["injs"] - out["dels"]).reshape(-1) == golden[b + "ohmic"]).all()
    assert (out["nr_part"] == golden[b + "nrpart_last"]).all()
    # one table set for every region gives another run
    dev.set_region_models([])
    other = dev.run(models[0], O.PM_CIC, acc, omega, dt, int(nsteps), int(seed), channel=CHANNEL)
    assert not (other["chan"] == golden[b + "chan"]).all()


def test_contact_rule_keeps_the_first_in_array_order(dev):
    """Excess particles at a contact node are the last ones in array order;
    the deficit is injected (ceil of expected - kept)."""
    exp_node = dev.cell_volume * 1e25 * 0.5  # XMIN node (0, 3): about 5
    pos = np.array([[0.1e-9, 3.1e-9]] * 8 + [[10e-9, 3e-9]] + [[0.2e-9, 2.9e-9]] * 2)
    keep, inj, exp = dev.contact_rule(pos)
    node = 3 * 21 + 0
    assert abs(exp[node] - exp_node) < 1e-12
    # count while nrPart < expected: ceil(expected) kept (1e-24 * 1e25 * 0.5
    # rounds to just above 5, so six are kept)
    kept = int(np.ceil(exp[node]))
    assert list(keep) == [1] * kept + [0] * (8 - kept) + [1] + [0] * 2
    assert inj[node] == 0
    keep, inj, exp = dev.contact_rule(pos[:2])
    assert inj[node] == int(np.ceil(exp[node] - 2)) and list(keep) == [1, 1]


def _cols(text):
    return [[float(x) for x in line.split()] for line in text.strip().split("\n")]


def test_emc_simulation_results_from_the_oracle_run(dev, golden):
    """emcSimulation::execute with setPoissonInterval(2) and a channel
    region: transient steps are excluded from the currents, the contact
    current is the running mean of (left - (injected - deleted)) * charge /
    dt, the drift current is the mean over the non-transient steps, and the
    final potential average covers the last nrStepsForFinalAvg steps."""
    seed, t_sim, dt, t_tr, n_avg, interval = golden["sim.cfg"]
    n_steps, n_tr = int(round(t_sim / dt)), int(round(t_tr / dt))
    out = dev.run(_model(), O.PM_CIC, 1e-8, 1.8, dt, n_steps, int(seed), interval=int(interval),
                  channel=CHANNEL)
    chan = out["chan"][n_tr:]
    assert np.isclose(chan.mean(), golden["sim.avg_drift_current"][0], rtol=1e-12, atol=0)
    netto = (out["exits"] - (out["injs"] - out["dels"]))[n_tr:]
    current = np.cumsum(netto, axis=0) / np.arange(1, len(netto) + 1)[:, None] * (-O.Q / dt)
    rows = _cols(bytes(golden["simfile.sim_ElectronsCurrent"]).decode())
    assert len(rows) == n_steps - n_tr
    for s, r in enumerate(rows):
        assert np.isclose(r[0], t_tr + (s + 1) * dt, rtol=1e-5)
        assert r[1:3] == netto[s].tolist()
        assert np.allclose(r[3:5], current[s], rtol=1e-5, atol=0)
    lines = bytes(golden["simfile.sim_PotentialAvg"]).decode().strip().split("\n")
    nx, ny = (int(x) for x in lines[0].split())
    avg = np.array([[float(x) for x in l.split()] for l in lines[1:]])
    expect = out["pot"][n_steps - int(n_avg):].mean(axis=0).reshape(ny, nx) * dev.Vt
    assert avg.shape == (ny, nx)
    assert np.allclose(avg, expect, rtol=1e-5, atol=1e-9)
