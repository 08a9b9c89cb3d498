"""2-D device path on the GPU against the oracle and the reference's fixtures.

* charge assignment, force interpolation (NGP / CIC / NEC) and the E-field
  on the reference's known-answer positions and potentials: bit-exact (FP64);
  the same for the 3-D NGP and CIC schemes and the 3-D grid-point E-field
  (emcgpu_deposit3, emcgpu_interpolate3, emcgpu_efield3);
* Poisson: the GPU's Newton / FP32-multigrid solution of the SOR equations
  against the oracle's SOR fixed point (both converged far below the
  tolerance) and against the discrete equation itself, also with two gate
  (Robin) contacts against the reference's recorded SOR solutions;
* the device particle loop under the reference's frozen field with the
  reference's recorded per-particle draws (replay): exit step and contact
  and event counts exact, positions / k / E within 1e-12 relative;
* the ohmic-contact rule on positions in array order: kept flags and
  injection counts exact;
* a self-consistent run (Poisson every step, contacts, injection, charge
  assignment): conservation of the ensemble counters, the FP64 check sum,
  Poisson convergence, graph launch equal to eager launch;
* the same loop with two gates (reflecting, Robin terms in every solve), by
  dense corrections and by multigrid;
* the mosfet2D example's 126 x 101 grid, whose multigrid data exceed the LDS
  (the solver keeps them in a global workspace): Poisson against the SOR
  fixed point and the discrete equations, and a self-consistent run against
  the oracle's twin;
* Poisson every third step and the drift-velocity channel current (over the
  middle of the device, and over all of it, where deleted and injected
  particles count) against the oracle's Philox twin; the running averages
  of potential and concentration against the per-step fields;
* the C++ host API's emcSimulation (resistor2DGpu example): the result
  files of the reference's layout, its equilibrium potential against the
  reference's, its contact currents against the per-step tallies.
"""
import numpy as np
import pytest

from emc_models import (DEVICE_GATED, DEVICE_SMALL, MODELS_DEVICE, device_region_specs,
                        mosfet_device, resistor_device)
from oracle import oracle_py as O

pytestmark = pytest.mark.gpu

REL_TOL = 1e-12


def _emc():
    from viennaemc_amd import emcgpu
    return emcgpu


def _ctx(dev=DEVICE_SMALL, pm=None, precision=0, tol=1e-13, max_iter=60):
    emcgpu = _emc()
    ctx = emcgpu.Context(precision=precision)
    ctx.setup(0, MODELS_DEVICE["device_para"])
    ctx.device_setup(dev, pm=emcgpu.PM_CIC if pm is None else pm, poisson_tol=tol,
                     poisson_max_iter=max_iter)
    return ctx


@pytest.fixture(scope="module")
def odev():
    return O.OracleDevice(DEVICE_SMALL)


@pytest.mark.parametrize("name", ["ngp", "cic", "nec"])
def test_pm_known_answers_bit_exact(golden, name):
    emcgpu = _emc()
    pm = {"ngp": emcgpu.PM_NGP, "cic": emcgpu.PM_CIC, "nec": emcgpu.PM_NEC}[name]
    print("PM known answers", name, flush=True)
    ctx = _ctx(pm=pm)
    pos = golden["device.kat.pos"].reshape(-1, 2)
    # every particle's weights are the reference's bit for bit ...
    od = O.OracleDevice(DEVICE_SMALL)
    opm = {"ngp": O.PM_NGP, "cic": O.PM_CIC, "nec": O.PM_NEC}[name]
    for p in pos[:40]:
        assert (ctx.deposit(p[None, :]) == od.assign(opm, p[None, :])).all()
    # ... and the grid sums them in atomic (unordered) order
    g = ctx.deposit(pos)
    ref = golden["device.kat.assign_" + name]
    assert np.allclose(g, ref, rtol=1e-15, atol=1e-15)
    assert abs(g.sum() - pos.shape[0]) < 1e-12
    f = ctx.interpolate(golden["device.kat.ex"], golden["device.kat.ey"], pos)
    ref = golden["device.kat.force_" + name].reshape(-1, 3)
    assert (f == ref[:, :2]).all()
    ctx.close()


def test_efield_bit_exact(golden):
    emcgpu = _emc()
    print("E-field known answers", flush=True)
    pot = golden["device.kat.pot"]
    for pm, tag in ((emcgpu.PM_CIC, "grid"), (emcgpu.PM_NEC, "edge")):
        ctx = _ctx(pm=pm)
        ex, ey = ctx.efield(pot)
        assert (ex == golden["device.kat.efield_%s_x" % tag]).all()
        assert (ey == golden["device.kat.efield_%s_y" % tag]).all()
        ctx.close()


@pytest.mark.parametrize("name", ["ngp", "cic"])
def test_pm3_known_answers_bit_exact(golden, name):
    """3-D assignToMesh / interpolateForce on the reference's 96 positions of
    an 8 x 6 x 5 grid (spacing 1, 0.5, 1 nm): each particle's weights bit
    for bit against the oracle, the grid (atomic order) against the
    reference's within 1e-15, the forces bit-exact."""
    emcgpu = _emc()
    pm = {"ngp": emcgpu.PM_NGP, "cic": emcgpu.PM_CIC}[name]
    opm = {"ngp": O.PM_NGP, "cic": O.PM_CIC}[name]
    print("3-D PM known answers", name, flush=True)
    ext, h = golden["pm3.ext"].astype(np.int32), golden["pm3.spacing"]
    pos = golden["pm3.pos"].reshape(-1, 3)
    ctx = emcgpu.Context(precision=emcgpu.FP32)  # the operators are FP64 in any context
    for p in pos[:40]:
        assert (ctx.deposit3(ext, h, p[None, :], pm=pm) == O.pm3_assign(opm, ext, h, p[None, :])).all()
    g = ctx.deposit3(ext, h, pos, pm=pm)
    assert np.allclose(g, golden["pm3.assign_" + name], rtol=1e-15, atol=1e-15)
    assert abs(g.sum() - pos.shape[0]) < 1e-12
    f = ctx.interpolate3(ext, h, golden["pm3.ex"], golden["pm3.ey"], golden["pm3.ez"], pos,
                         charge=-O.Q, pm=pm)
    assert (f.reshape(-1) == golden["pm3.force_" + name]).all()
    ctx.close()


def test_pm3_efield_bit_exact(golden):
    """3-D calcEFieldAtGridPts with the faces' boundary rule (contacts over
    parts of XMIN and ZMAX, as the reference's emcSurface reports them)."""
    emcgpu = _emc()
    print("3-D E-field known answers", flush=True)
    ext, h = golden["pm3.ext"].astype(np.int32), golden["pm3.spacing"]
    ctx = emcgpu.Context(precision=emcgpu.FP64)
    E = ctx.efield3(ext, h, float(golden["pm3.vt"][0]), golden["pm3.bc"].astype(np.int8),
                    golden["pm3.pot"])
    for d, c in enumerate("xyz"):
        assert (E[d] == golden["pm3.efield_" + c]).all()
    ctx.close()


def _residual(dev, phi, n, C, eq, spec=DEVICE_SMALL):
    """max |F| of the discrete SOR equations at the non-ohmic nodes (with
    the Robin term g (V_g - phi) at the nodes of a gate in spec)"""
    nx, ny = dev.ext
    gterm = np.zeros((ny, nx))
    gdiag = np.zeros((ny, nx))
    for ct in spec["contacts"]:
        if len(ct) != 7:
            continue
        b, v, lo, hi, eps_ox, t_ox, barrier = ct
        gf = 2 * (eps_ox / spec["eps_r"]) / (t_ox / dev.debye) * (1e-9 / dev.debye)
        vg = (barrier + (0.0 if eq else v)) / dev.Vt
        iy = ny - 1 if b == 3 else 0
        assert b in (2, 3)
        for ix in range(int(round(lo / 1e-9)), int(round(hi / 1e-9)) + 1):
            gterm[iy, ix] += gf * vg
            gdiag[iy, ix] += gf
    phi = phi.reshape(ny, nx)
    h = 1e-9 / dev.debye
    ohm = dev.grids()["contact"].reshape(ny, nx) >= 0
    worst = 0.0
    for iy in range(ny):
        for ix in range(nx):
            if ohm[iy, ix]:
                continue
            xm = phi[iy, ix - 1] if ix > 0 else phi[iy, ix + 1]
            xp = phi[iy, ix + 1] if ix < nx - 1 else phi[iy, ix - 1]
            ym = phi[iy - 1, ix] if iy > 0 else phi[iy + 1, ix]
            yp = phi[iy + 1, ix] if iy < ny - 1 else phi[iy - 1, ix]
            nn = np.exp(phi[iy, ix]) if eq else n.reshape(ny, nx)[iy, ix]
            r = (xm + xp + ym + yp - 4 * phi[iy, ix]) + h * h * (
                np.exp(-phi[iy, ix]) - nn + C.reshape(ny, nx)[iy, ix])
            r += gterm[iy, ix] - gdiag[iy, ix] * phi[iy, ix]
            worst = max(worst, abs(r))
    return worst


def test_poisson_matches_sor_fixed_point(odev, golden):
    print("Poisson vs SOR", flush=True)
    ctx = _ctx(tol=1e-14)
    pot0 = odev.init_potential()
    ref_eq, _ = odev.sor(pot0, accuracy=1e-14, omega=1.8)
    eq, it, res = ctx.poisson(pot0, equilibrium=True, reset_bc=True)
    print("equilibrium: %d Newton iterations, residual %.2e, max |d| %.2e" % (
        it, res, np.abs(eq - ref_eq).max()), flush=True)
    assert res < 1e-13
    assert np.abs(eq - ref_eq).max() < 1e-9
    conc = golden["device.kat.sor_conc"]
    ref_neq, _ = odev.sor(ref_eq, conc=conc, accuracy=1e-14, omega=1.8)
    neq, it2, res2 = ctx.poisson(eq, conc=conc, equilibrium=False, reset_bc=True)
    print("non-equilibrium: %d iterations, residual %.2e, max |d| %.2e" % (
        it2, res2, np.abs(neq - ref_neq).max()), flush=True)
    assert res2 < 1e-13
    assert np.abs(neq - ref_neq).max() < 1e-9
    C = odev.grids()["doping"] / DEVICE_SMALL["ni"]
    assert _residual(odev, neq, conc, C, eq=False) < 1e-12
    ctx.close()


def test_poisson_with_gates_matches_sor(golden):
    """Two gate contacts (addGateContact on YMAX and YMIN): the GPU solution
    against the oracle's SOR at 1e-14 and the reference's own SOR solutions
    (1e-12 stopping rule), equilibrium (barrier height only) and
    non-equilibrium (barrier plus gate voltage)."""
    print("Poisson with gates vs SOR", flush=True)
    od = O.OracleDevice(DEVICE_GATED)
    ctx = _ctx(dev=DEVICE_GATED, tol=1e-14)
    pot0 = od.init_potential()
    ref_eq, _ = od.sor(pot0, accuracy=1e-14, omega=1.8)
    eq, it, res = ctx.poisson(pot0, equilibrium=True, reset_bc=True)
    print("equilibrium: %d iterations, residual %.2e, max |d| %.2e, vs reference %.2e" % (
        it, res, np.abs(eq - ref_eq).max(), np.abs(eq - golden["device.gate.sor_eq"]).max()),
        flush=True)
    assert res < 1e-13
    assert np.abs(eq - ref_eq).max() < 1e-9
    assert np.abs(eq - golden["device.gate.sor_eq"]).max() < 1e-8
    conc = golden["device.gate.sor_conc"]
    ref_neq, _ = od.sor(golden["device.gate.sor_eq"], conc=conc, accuracy=1e-14, omega=1.8)
    neq, it2, res2 = ctx.poisson(golden["device.gate.sor_eq"], conc=conc, equilibrium=False,
                                 reset_bc=True)
    print("non-equilibrium: %d iterations, residual %.2e, max |d| %.2e, vs reference %.2e" % (
        it2, res2, np.abs(neq - ref_neq).max(), np.abs(neq - golden["device.gate.sor_neq"]).max()),
        flush=True)
    assert res2 < 1e-13
    assert np.abs(neq - ref_neq).max() < 1e-9
    assert np.abs(neq - golden["device.gate.sor_neq"]).max() < 1e-8
    C = od.grids()["doping"] / DEVICE_GATED["ni"]
    assert _residual(od, neq, conc, C, eq=False, spec=DEVICE_GATED) < 1e-12
    assert _residual(od, eq, None, C, eq=True, spec=DEVICE_GATED) < 1e-12
    # the gate nodes are no Dirichlet nodes: the gate pulls them far from
    # the ungated solution
    assert np.abs(eq - golden["device.kat.sor_eq"]).max() > 1.0
    ctx.close()


def test_poisson_resistor_grid_converges():
    """200 x 50 nodes: the multigrid Newton solve reaches 1e-10 in a few
    iterations from the charge-neutral guess and from a warm start."""
    print("Poisson 200x50", flush=True)
    dev = resistor_device()
    ctx = _ctx(dev=dev, tol=1e-10)
    od = O.OracleDevice(dev)
    pot0 = od.init_potential()
    eq, it, res = ctx.poisson(pot0, equilibrium=True, reset_bc=True)
    print("equilibrium: %d iterations, residual %.2e" % (it, res), flush=True)
    assert res <= 1e-10 and it < 60
    n = np.exp(eq)  # electrons in equilibrium: a bias then moves the potential
    neq, it2, res2 = ctx.poisson(eq, conc=n, equilibrium=False, reset_bc=True)
    print("biased: %d iterations, residual %.2e" % (it2, res2), flush=True)
    assert res2 <= 1e-10 and it2 < 30
    ref, _ = od.sor(eq, conc=n, accuracy=1e-11, omega=1.9)
    print("max |phi - phi_SOR| = %.2e" % np.abs(neq - ref).max(), flush=True)
    assert np.abs(neq - ref).max() < 1e-5
    again, it3, res3 = ctx.poisson(neq, conc=n, equilibrium=False, reset_bc=False)
    assert it3 <= 1 and res3 <= 1e-10
    ctx.close()


def test_device_trajectories_replay(golden):
    emcgpu = _emc()
    b = "device.traj."
    np_, ns, seed = [int(x) for x in golden[b + "meta"]]
    dt = golden[b + "dt"][0]
    print("device replay:", np_, "particles", ns, "steps", flush=True)
    init = golden[b + "init"].reshape(np_, 8)  # x, y, k, E, tau, grainTau
    ctx = _ctx()
    zero = np.zeros(np_, dtype=np.int32)
    ctx.device_set_particles(init[:, 0:2], init[:, 2:5], init[:, 5], init[:, 6], zero, zero)
    ctx.device_set_efield(golden[b + "ex"], golden[b + "ey"])
    draws = O.mt64_particle_draws(seed, np_, 512)
    cursor = golden[b + "init_draws"].astype(np.uint64).copy()
    grain = init[:, 7].copy()
    tr = ctx.device_drift(dt, ns, np_, mode=emcgpu.REPLAY, draws=draws, cursor=cursor,
                          grain_tau=grain, trace=True)
    rem = tr["exit_step"]
    assert (rem == golden[b + "removed_step"]).all()
    who = golden[b + "removed_who"]
    assert (tr["exit_contact"][who] == golden[b + "removed_contact"]).all()
    assert (tr["events"] == golden[b + "nevents"].reshape(ns, np_)).all()
    pos_ref = golden[b + "pos"].reshape(ns, np_, 2)
    rel = np.abs(tr["pos"] - pos_ref).max() / np.abs(pos_ref).max()
    print("positions: max rel %.2e" % rel, flush=True)
    assert rel < REL_TOL
    snaps = golden[b + "snap_steps"]
    k = tr["k"][snaps].reshape(-1, 3)
    kr = golden[b + "k"].reshape(-1, 3)
    assert (np.linalg.norm(k - kr, axis=1) / np.linalg.norm(kr, axis=1)).max() < REL_TOL
    E = tr["E"][snaps].reshape(-1)
    assert (np.abs(E - golden[b + "E"]) / golden[b + "E"]).max() < REL_TOL
    total = golden[b + "init_draws"] + golden[b + "draws"].reshape(ns, np_).sum(axis=0)
    assert (cursor == total).all()
    ctx.close()


def test_contact_rule_exact(odev):
    print("contact rule", flush=True)
    rng = np.random.default_rng(7)
    # 400 particles, half of them inside the ohmic columns (x < 0.5 nm or
    # x > 19.5 nm), in random array order
    n = 400
    x = np.where(rng.random(n) < 0.5, rng.random(n) * 0.5e-9,
                 rng.random(n) * 20e-9)
    x = np.where(rng.random(n) < 0.3, 20e-9 - rng.random(n) * 0.5e-9, x)
    pos = np.stack([x, rng.random(n) * 6e-9], axis=1)
    ctx = _ctx()
    keep, inj = ctx.contacts(pos)
    okeep, oinj, _ = odev.contact_rule(pos)
    assert (keep == okeep).all()
    assert (inj == oinj).all()
    assert (keep == 0).sum() > 10
    # few particles: everything kept, deficits injected
    keep, inj = ctx.contacts(pos[:20])
    okeep, oinj, _ = odev.contact_rule(pos[:20])
    assert (keep == okeep).all() and (inj == oinj).all()
    assert keep.all() and inj.sum() > 0
    ctx.close()


@pytest.mark.parametrize("precision", [0, 1])
def test_self_consistent_run(precision):
    """Resistor at 0.5 V: counters conserve the ensemble, the charge check
    sum stays at rounding level, every step's Poisson solve converges."""
    dev = resistor_device(n_particles=200_000)
    print("self-consistent run, precision", precision, flush=True)
    ctx = _ctx(dev=dev, precision=precision, tol=1e-10)
    n0 = ctx.device_init(1234)
    assert abs(n0 - 200_000) < 2000
    out = ctx.device_step(0.1e-15, 200, seed=1234)
    n = out["n_particles"]
    d = (out["injected"] - out["exits"] - out["deleted"]).sum(axis=1)
    assert n[0] == n0 + d[0] and (np.diff(n) == d[1:]).all()
    assert (out["poisson_resid"] <= 1e-10).all()
    cs = out["checksum"].max()
    print("n %d -> %d, exits %d, injected %d, deleted %d, Newton its %.2f, check sum %.2e" % (
        n0, n[-1], out["exits"].sum(), out["injected"].sum(), out["deleted"].sum(),
        out["poisson_iters"].mean(), cs), flush=True)
    assert cs < (1e-12 if precision == 0 else 1e-5)
    assert out["exits"].sum() > 0 and out["injected"].sum() > 0
    p = ctx.device_get_particles()
    assert p["E"].size == n[-1]
    assert (p["pos"][:, 0] >= 0).all() and (p["pos"][:, 0] <= dev["max_pos"][0]).all()
    f = ctx.device_fields()
    assert abs(f["nr_part"].sum() - n[-1]) < (1e-6 if precision == 0 else 1.0)
    ctx.close()


@pytest.mark.parametrize("precision", [0, 1])
def test_dense_poisson_corrections_match_multigrid(precision):
    """The per-step solves with Newton corrections from an inverse of the
    discrete operator against the multigrid path alone (POISSON_MULTIGRID)
    on the 0.5 V resistor from the same seed.  The resistor's Dirichlet
    nodes are whole columns, so POISSON_AUTO takes the separable inverse;
    POISSON_DENSE takes the dense one.  All meet the stopping rule at every
    step, the corrections leave the multigrid solve (almost) nothing to do
    (the exact separable inverse needs one correction per step), and the
    potentials and ensembles agree to the solver tolerance."""
    emcgpu = _emc()
    dev = resistor_device(n_particles=200_000)
    print("inverse corrections vs multigrid, precision", precision, flush=True)
    res = {}
    for mode in (emcgpu.POISSON_MULTIGRID, emcgpu.POISSON_AUTO, emcgpu.POISSON_DENSE):
        ctx = _ctx(dev=dev, precision=precision, tol=1e-10)
        ctx.device_init(777)
        out = ctx.device_step(0.1e-15, 40, seed=777, poisson_mode=mode)
        res[mode] = (out, ctx.device_fields()["pot"], ctx.device_get_particles())
        ctx.close()
    mg = res[emcgpu.POISSON_MULTIGRID]
    for out, _, _ in res.values():
        assert (out["poisson_resid"] <= 1e-10).all()
    # the first step sets the contacts' voltages and solves by multigrid;
    # afterwards the recorded work is corrections + V-cycles
    work_mg = mg[0]["poisson_iters"][1:]
    scale = np.abs(mg[1]).max()
    for mode, name, most in ((emcgpu.POISSON_AUTO, "separable", 1),
                             (emcgpu.POISSON_DENSE, "dense", 2)):
        out, pot, _ = res[mode]
        work = out["poisson_iters"][1:]
        print("work per step: multigrid %.2f V-cycles, %s %.2f (max %d), residual <= %.1e"
              % (work_mg.mean(), name, work.mean(), work.max(), out["poisson_resid"][1:].max()),
              flush=True)
        assert (work <= most).mean() >= 0.9 and work.mean() < 0.5 * work_mg.mean()
        # the solves agree to their tolerance; over 40 steps FP32 positions
        # turn that into differences of ~1e-8 of the scale
        dpot = np.abs(pot - mg[1]).max() / scale
        print("  max |pot - multigrid pot| = %.2e of the scale" % dpot, flush=True)
        assert dpot < (1e-8 if precision == 0 else 1e-7)
        assert abs(int(out["n_particles"][-1]) - int(mg[0]["n_particles"][-1])) <= 2


def test_graph_launch_equals_eager():
    dev = resistor_device(n_particles=50_000)
    print("device graph vs eager", flush=True)
    res = {}
    for graph in (False, True):
        ctx = _ctx(dev=dev, precision=0, tol=1e-10)
        ctx.device_init(99)
        a = ctx.device_step(0.1e-15, 70, seed=99, graph=graph)
        b = ctx.device_step(0.1e-15, 64, seed=99, graph=graph)
        p = ctx.device_get_particles()
        res[graph] = (a, b, p)
        ctx.close()
    for x, y in zip(res[False][:2], res[True][:2]):
        for key in ("exits", "deleted", "injected", "n_particles"):
            assert (x[key] == y[key]).all(), key
    pa, pb = res[False][2], res[True][2]
    assert pa["E"].size == pb["E"].size
    # the deposit sums in a fixed order except for the few FP64 atomic
    # corrections; trajectories agree to rounding
    assert np.abs(pa["pos"] - pb["pos"]).max() < 1e-15


def test_self_consistent_run_matches_oracle_twin():
    """The whole self-consistent loop (equilibrium, initial particles,
    Poisson every step, drift/scatter with the reference's boundaries,
    contact equilibrium in array order, injection, CIC assignment) against
    the oracle's Philox twin: ensemble counts per step and per contact
    exact, final ensemble in array order within 1e-9 (the two Poisson
    solvers agree to their tolerance, not bit for bit)."""
    emcgpu = _emc()
    dev = dict(DEVICE_SMALL, width=1e-5)
    seed, ns, dt = 4242, 60, 1e-15
    od = O.OracleDevice(dev)
    M = O.OracleModel(MODELS_DEVICE["device_para"])
    ref = od.run_philox(M, O.PM_CIC, 1e-14, 1.8, dt, ns, seed, O.thermal_voltage(300.0),
                        max_state=100000)
    print("twin: %d particles, %d steps" % (ref["n_init"][0], ns), flush=True)
    ctx = _ctx(dev=dev, tol=1e-14, max_iter=80)
    n0 = ctx.device_init(seed)
    assert n0 == ref["n_init"][0]
    eq = ctx.device_fields()["pot"]
    assert np.abs(eq - ref["eq_pot"]).max() < 1e-9
    out = ctx.device_step(dt, ns, seed=seed)
    assert (out["n_particles"] == ref["n_part"]).all()
    assert (out["exits"] == ref["exits"]).all()
    assert (out["deleted"] == ref["dels"]).all()
    assert (out["injected"] == ref["injs"]).all()
    p = ctx.device_get_particles()
    st = ref["state"]
    assert p["E"].size == st.shape[0]
    assert (p["valley"] == st[:, 7]).all() and (p["sub"] == st[:, 8]).all()
    dpos = np.abs(p["pos"] - st[:, 0:2]).max() / 20e-9
    dk = (np.linalg.norm(p["k"] - st[:, 2:5], axis=1) / np.linalg.norm(st[:, 2:5], axis=1)).max()
    dE = (np.abs(p["E"] - st[:, 5]) / st[:, 5]).max()
    print("final ensemble: |dpos| %.2e, |dk| %.2e, |dE| %.2e" % (dpos, dk, dE), flush=True)
    assert dpos < 1e-9 and dk < 1e-9 and dE < 1e-9
    f = ctx.device_fields()
    assert np.allclose(f["nr_part"], ref["nr_part"], rtol=1e-9, atol=1e-9)
    ctx.close()


# a gate over the whole YMAX boundary keeps the device separable
DEVICE_GATED_FULL = dict(DEVICE_SMALL, contacts=DEVICE_SMALL["contacts"] + [
    (3, 0.3, 0.0, 20e-9, 3.9, 1e-9, 0.3)])


@pytest.mark.parametrize("gates,mode", [("partial", "auto"), ("partial", "multigrid"),
                                        ("full", "auto")])
def test_gated_run_matches_oracle_twin(gates, mode):
    """The self-consistent loop with gates: particles reaching a gate are
    reflected (only the ohmic contacts delete and inject), every solve
    carries the gates' Robin terms.  Against the oracle's Philox twin as
    above: two partial gates by the dense corrections (auto) and by
    multigrid alone, a gate over a whole boundary by the separable
    corrections (auto)."""
    emcgpu = _emc()
    dev = dict(DEVICE_GATED if gates == "partial" else DEVICE_GATED_FULL, width=1e-5)
    seed, ns, dt = 777, 50, 1e-15
    od = O.OracleDevice(dev)
    M = O.OracleModel(MODELS_DEVICE["device_para"])
    ref = od.run_philox(M, O.PM_CIC, 1e-14, 1.8, dt, ns, seed, O.thermal_voltage(300.0),
                        max_state=100000)
    print("gated twin (%s gates, %s): %d particles, %d steps" % (gates, mode, ref["n_init"][0], ns),
          flush=True)
    ctx = _ctx(dev=dev, tol=1e-14, max_iter=80)
    assert ctx.device_init(seed) == ref["n_init"][0]
    assert np.abs(ctx.device_fields()["pot"] - ref["eq_pot"]).max() < 1e-9
    pmode = emcgpu.POISSON_AUTO if mode == "auto" else emcgpu.POISSON_MULTIGRID
    out = ctx.device_step(dt, ns, seed=seed, poisson_mode=pmode)
    assert (out["n_particles"] == ref["n_part"]).all()
    assert (out["exits"] == ref["exits"]).all()
    assert (out["deleted"] == ref["dels"]).all()
    assert (out["injected"] == ref["injs"]).all()
    # 1e-14 is at the FP64 rounding floor of the residual: every solve ends
    # within a few units of it
    assert (out["poisson_resid"] <= 1e-13).all()
    p = ctx.device_get_particles()
    st = ref["state"]
    assert p["E"].size == st.shape[0]
    dpos = np.abs(p["pos"] - st[:, 0:2]).max() / 20e-9
    dE = (np.abs(p["E"] - st[:, 5]) / st[:, 5]).max()
    print("final ensemble: |dpos| %.2e, |dE| %.2e, Poisson work %s" % (
        dpos, dE, np.bincount(out["poisson_iters"])), flush=True)
    assert dpos < 1e-9 and dE < 1e-9
    # the potential of the last step's solve
    assert np.abs(ctx.device_fields()["pot"] - ref["pot"][-1]).max() < 1e-9
    ctx.close()


def test_poisson_beyond_lds_matches_sor():
    """The mosfet2D example's 126 x 101 grid (p-type bulk, n+ source and
    drain, a gate between them): its multigrid levels exceed the LDS, so
    the solver keeps them in a global workspace.  Equilibrium and biased
    solutions against the oracle's SOR at 1e-14 and against the discrete
    equations with the gate's Robin term."""
    print("Poisson beyond the LDS (mosfet2D grid)", flush=True)
    dev = mosfet_device()
    od = O.OracleDevice(dev)
    assert tuple(od.ext) == (126, 101)
    ctx = _ctx(dev=dev, tol=1e-14, max_iter=200)
    pot0 = od.init_potential()
    ref_eq, _ = od.sor(pot0, accuracy=1e-14, omega=1.8)
    eq, it, res = ctx.poisson(pot0, equilibrium=True, reset_bc=True)
    print("equilibrium: %d iterations, residual %.2e, max |d| %.2e" % (
        it, res, np.abs(eq - ref_eq).max()), flush=True)
    assert res < 1e-13
    assert np.abs(eq - ref_eq).max() < 1e-9
    conc = np.exp(ref_eq)  # equilibrium electrons under the applied voltages
    ref_neq, _ = od.sor(ref_eq, conc=conc, accuracy=1e-14, omega=1.8)
    neq, it2, res2 = ctx.poisson(ref_eq, conc=conc, equilibrium=False, reset_bc=True)
    print("biased: %d iterations, residual %.2e, max |d| %.2e" % (
        it2, res2, np.abs(neq - ref_neq).max()), flush=True)
    assert res2 < 1e-13
    assert np.abs(neq - ref_neq).max() < 1e-9
    C = od.grids()["doping"] / dev["ni"]
    assert _residual(od, eq, None, C, eq=True, spec=dev) < 1e-12
    assert _residual(od, neq, conc, C, eq=False, spec=dev) < 1e-12
    # the drain and gate voltages move the potential
    assert np.abs(neq - eq).max() > 10.0
    ctx.close()


@pytest.fixture(scope="module")
def mosfet_twin():
    """The oracle's Philox twin of the mosfet2D device at a width of 20 nm
    (about 3000 electrons), 6 steps of 1 fs."""
    dev = mosfet_device(width=2e-8)
    seed, ns, dt = 515, 6, 1e-15
    od = O.OracleDevice(dev)
    M = O.OracleModel(MODELS_DEVICE["device_para"])
    ref = od.run_philox(M, O.PM_CIC, 1e-14, 1.8, dt, ns, seed, O.thermal_voltage(300.0),
                        max_state=100000)
    return dev, seed, ns, dt, ref


@pytest.mark.parametrize("mode", ["auto", "multigrid"])
def test_mosfet_run_matches_oracle_twin(mosfet_twin, mode):
    """The self-consistent loop on the mosfet2D grid (multigrid data in the
    global workspace; auto: dense Newton corrections, then the multigrid
    check) against the oracle's Philox twin: ensemble counts per step and
    per contact exact, final ensemble and potential within 1e-9."""
    emcgpu = _emc()
    dev, seed, ns, dt, ref = mosfet_twin
    print("mosfet twin (%s): %d particles, %d steps, exits %s" % (
        mode, ref["n_init"][0], ns, ref["exits"].sum(axis=0)), flush=True)
    ctx = _ctx(dev=dev, tol=1e-14, max_iter=200)
    assert ctx.device_init(seed) == ref["n_init"][0]
    assert np.abs(ctx.device_fields()["pot"] - ref["eq_pot"]).max() < 1e-9
    pmode = emcgpu.POISSON_AUTO if mode == "auto" else emcgpu.POISSON_MULTIGRID
    out = ctx.device_step(dt, ns, seed=seed, poisson_mode=pmode)
    assert (out["n_particles"] == ref["n_part"]).all()
    assert (out["exits"] == ref["exits"]).all()
    assert (out["deleted"] == ref["dels"]).all()
    assert (out["injected"] == ref["injs"]).all()
    assert (out["poisson_resid"] <= 1e-13).all()
    p = ctx.device_get_particles()
    st = ref["state"]
    assert p["E"].size == st.shape[0]
    assert (p["valley"] == st[:, 7]).all() and (p["sub"] == st[:, 8]).all()
    dpos = np.abs(p["pos"] - st[:, 0:2]).max() / 125e-9
    dE = (np.abs(p["E"] - st[:, 5]) / st[:, 5]).max()
    print("final ensemble: |dpos| %.2e, |dE| %.2e, Poisson work %s" % (
        dpos, dE, np.bincount(out["poisson_iters"])), flush=True)
    assert dpos < 1e-9 and dE < 1e-9
    assert np.abs(ctx.device_fields()["pot"] - ref["pot"][-1]).max() < 1e-9
    ctx.close()


def test_region_tables_run_matches_oracle_twin():
    """Mechanisms registered for some doping regions only (region_tables:
    instance r holds the tables of region r): the loop against the oracle's
    Philox twin with the same per-region tables, counts per step exact and
    the final ensemble within 1e-9; instance 0's tables everywhere give
    another run."""
    emcgpu = _emc()
    dev = dict(DEVICE_SMALL, width=1e-5)
    seed, ns, dt = 2468, 50, 1e-15
    specs = device_region_specs()
    od = O.OracleDevice(dev)
    models = [O.OracleModel(sp) for sp in specs]
    od.set_region_models(models)
    ref = od.run_philox(models[0], O.PM_CIC, 1e-14, 1.8, dt, ns, seed, O.thermal_voltage(300.0),
                        max_state=100000)
    print("region-table twin: %d particles, %d steps" % (ref["n_init"][0], ns), flush=True)
    res = {}
    for per_region in (True, False):
        ctx = emcgpu.Context(precision=0, n_instances=3)
        for r, sp in enumerate(specs):
            ctx.setup(r, sp)
        ctx.device_setup(dev, pm=emcgpu.PM_CIC, poisson_tol=1e-14, poisson_max_iter=80,
                         region_tables=per_region)
        assert ctx.device_init(seed) == ref["n_init"][0]
        out = ctx.device_step(dt, ns, seed=seed)
        res[per_region] = (out, ctx.device_get_particles())
        ctx.close()
    out, p = res[True]
    assert (out["n_particles"] == ref["n_part"]).all()
    assert (out["exits"] == ref["exits"]).all()
    assert (out["deleted"] == ref["dels"]).all()
    assert (out["injected"] == ref["injs"]).all()
    st = ref["state"]
    assert p["E"].size == st.shape[0]
    dpos = np.abs(p["pos"] - st[:, 0:2]).max() / 20e-9
    dE = (np.abs(p["E"] - st[:, 5]) / st[:, 5]).max()
    print("final ensemble: |dpos| %.2e, |dE| %.2e" % (dpos, dE), flush=True)
    assert dpos < 1e-9 and dE < 1e-9
    other = res[False][1]
    assert other["E"].size != st.shape[0] or np.abs(other["E"] - st[:, 5]).max() > 1e-6


@pytest.mark.parametrize("region", ["middle", "whole"])
def test_poisson_interval_and_channel_current_match_oracle_twin(region):
    """setPoissonInterval(3) + setChannelCurrentRegion against the oracle's
    twin: counts per step and contact exact, solves only on every third
    step, channel current per step within 1e-6 of the twin's (the particles
    agree to ~1e-10, the sums cancel in part).  Over the whole device the
    particles deleted or injected at the contacts are inside the region, so
    the current includes the step kernel's corrections for them."""
    emcgpu = _emc()
    dev = dict(DEVICE_SMALL, width=1e-5)
    seed, ns, dt, every = 5150, 45, 1e-15, 3
    chan = (5e-9, 15e-9, 10e-9) if region == "middle" else (0.0, 20e-9, 20e-9)
    od = O.OracleDevice(dev)
    M = O.OracleModel(MODELS_DEVICE["device_para"])
    ref = od.run_philox(M, O.PM_CIC, 1e-14, 1.8, dt, ns, seed, O.thermal_voltage(300.0),
                        interval=every, channel=chan)
    print("interval / channel twin (%s): %d particles" % (region, ref["n_init"][0]), flush=True)
    ctx = _ctx(dev=dev, tol=1e-14, max_iter=80)
    assert ctx.device_init(seed) == ref["n_init"][0]
    out = ctx.device_step(dt, ns, seed=seed, poisson_interval=every, channel=chan)
    assert (out["n_particles"] == ref["n_part"]).all()
    assert (out["exits"] == ref["exits"]).all()
    assert (out["deleted"] == ref["dels"]).all()
    assert (out["injected"] == ref["injs"]).all()
    solved = out["poisson_iters"] > 0
    assert (solved == (np.arange(ns) % every == 0)).all()
    if region == "whole":
        assert out["deleted"].sum() + out["injected"].sum() > 50
    c, r = out["channel_current"], ref["chan"]
    print("channel current: max |gpu - twin| / max |twin| = %.2e"
          % (np.abs(c - r).max() / np.abs(r).max()), flush=True)
    assert np.allclose(c, r, rtol=1e-6, atol=1e-6 * np.abs(r).max())
    ctx.close()


def test_running_averages_equal_the_mean_of_the_step_fields():
    """updateAverageCharacteristics on the GPU: after steps run with
    accumulate, the averages are the mean of the potential and the
    concentration each of those steps left (same summation order:
    bit-exact); a reset clears them."""
    emcgpu = _emc()
    ctx = _ctx(tol=1e-12)
    ctx.device_init(99)
    ctx.device_step(1e-15, 20, seed=99, poisson_interval=2)
    pot_sum, conc_sum = np.zeros(ctx.nn), np.zeros(ctx.nn)
    n = 12
    for _ in range(n):
        ctx.device_step(1e-15, 1, seed=99, poisson_interval=2, accumulate=True)
        f = ctx.device_fields()
        pot_sum += f["pot"]
        conc_sum += f["conc"]
    pot, conc, k = ctx.device_averages(reset=True)
    assert k == n
    assert (pot == pot_sum / n).all() and (conc == conc_sum / n).all()
    assert ctx.device_averages()[2] == 0
    ctx.close()


def _grid(text):
    lines = text.strip().split("\n")
    nx, ny = (int(x) for x in lines[0].split())
    g = np.array([[float(x) for x in l.split()] for l in lines[1:]])
    assert g.shape == (ny, nx)
    return g


def test_cpp_simulation_writes_the_reference_results(golden, tmp_path):
    """resistor2DGpu (small) is the harness's emcSimulation run (same device,
    times, Poisson interval, channel, averaging) on the GPU: it writes the
    same nine files with the same layout; its equilibrium potential and
    field match the reference's files to their printed digits; its contact
    current file is the running mean of its own per-step tallies, which the
    Python binding reproduces on the same seed."""
    import os
    import subprocess
    emcgpu = _emc()
    exe = os.path.join(os.path.dirname(emcgpu.__file__), "examples", "resistor2DGpu")
    seed = 3131
    prefix = str(tmp_path / "sim_")
    print("cpp device example", flush=True)
    txt = subprocess.check_output([exe, prefix, "small", "fp64", str(seed)],
                                  env=dict(os.environ, EMCGPU_THROW="1"), timeout=120).decode()
    names = sorted(k[len("simfile."):] for k in golden.files if k.startswith("simfile."))
    for name in names:
        ours = (tmp_path / (name + ".txt")).read_text()
        theirs = bytes(golden["simfile." + name]).decode()
        lo, lt = ours.split("\n"), theirs.split("\n")
        if name.endswith(("Eq", "Final")) and name.startswith("sim_Electrons") \
                and "Conc" not in name:
            # particle dumps: the ensemble sizes differ with the random
            # numbers; header, columns and numbering as the reference's
            assert lo[0] == lt[0], name
            assert abs(len(lo) - len(lt)) < 0.05 * len(lt), name
            for i, line in enumerate(lo[1:]):
                t = line.split()
                assert len(t) == 10 and int(t[0]) == i, (name, i)
            assert not ours.endswith("\n"), name
            continue
        assert len(lo) == len(lt), name
        # header (extent), or the step time of the first current line
        first = lo[0], lt[0]
        if name.endswith("Current"):
            first = first[0].split()[0], first[1].split()[0]
        assert first[0] == first[1], name
    for name in ("sim_PotentialEq",):
        ours = _grid((tmp_path / (name + ".txt")).read_text())
        theirs = _grid(bytes(golden["simfile." + name]).decode())
        assert np.allclose(ours, theirs, rtol=2e-5, atol=1e-9), name
    # E-field: printed digits and the SOR's accuracy (1e-8 V over 1 nm)
    e_scale = np.abs(_grid(bytes(golden["simfile.sim_EFieldXEq"]).decode())).max()
    for name in ("sim_EFieldXEq", "sim_EFieldYEq"):
        ours = _grid((tmp_path / (name + ".txt")).read_text())
        theirs = _grid(bytes(golden["simfile." + name]).decode())
        assert np.allclose(ours, theirs, rtol=1e-4, atol=1e-4 * e_scale), name
    # the same run through the Python binding: contact tallies -> currents
    ctx = _ctx(tol=1e-10, max_iter=50)
    ctx.device_init(seed)
    out = ctx.device_step(1e-15, 30, seed=seed, poisson_interval=2)
    netto = (out["exits"] - (out["injected"] - out["deleted"]))[10:]
    cur = np.cumsum(netto, axis=0) / np.arange(1, 21)[:, None] * (-O.Q / 1e-15)
    rows = [[float(x) for x in l.split()]
            for l in (tmp_path / "sim_ElectronsCurrent.txt").read_text().strip().split("\n")]
    assert len(rows) == 20
    for s, r in enumerate(rows):
        assert r[1:3] == netto[s].tolist(), s
        assert np.allclose(r[3:5], cur[s], rtol=1e-5, atol=0), s
    assert "drift current" in txt
    ctx.close()


def test_cpp_simulation_with_region_tables(tmp_path):
    """resistor2DGpu with mechanisms registered for some doping regions only:
    the C++ host API gives each region its own tables (one instance per
    region), and its per-step contact tallies equal the Python binding's run
    with the same per-region tables."""
    import os
    import subprocess
    emcgpu = _emc()
    exe = os.path.join(os.path.dirname(emcgpu.__file__), "examples", "resistor2DGpu")
    seed = 4141
    prefix = str(tmp_path / "sim_")
    print("cpp device example with region tables", flush=True)
    subprocess.check_output([exe, prefix, "small", "fp64", str(seed), "regions"],
                            env=dict(os.environ, EMCGPU_THROW="1"), timeout=120)
    rows = [[float(x) for x in l.split()]
            for l in (tmp_path / "sim_ElectronsCurrent.txt").read_text().strip().split("\n")]
    nettos = {}
    for per_region in (True, False):
        ctx = emcgpu.Context(precision=emcgpu.FP64, n_instances=3)
        for r, sp in enumerate(device_region_specs()):
            ctx.setup(r, sp)
        ctx.device_setup(DEVICE_SMALL, pm=emcgpu.PM_CIC, poisson_tol=1e-10, poisson_max_iter=50,
                         region_tables=per_region)
        ctx.device_init(seed)
        out = ctx.device_step(1e-15, 30, seed=seed, poisson_interval=2)
        nettos[per_region] = (out["exits"] - (out["injected"] - out["deleted"]))[10:]
        ctx.close()
    assert len(rows) == 20
    for s_, r in enumerate(rows):
        assert r[1:3] == nettos[True][s_].tolist(), s_
    assert (nettos[True] != nettos[False]).any()


def test_initial_particles_from_the_equilibrium_potential_match_oracle_twin():
    """usePotentialForInit: the initial count per node is exp(phi_eq) * n_i
    times the cell volume (halved on boundaries), then the reference's
    remainder rule; against the oracle's twin with the same option, the
    initial ensemble and 20 steps of counts per contact are exact."""
    emcgpu = _emc()
    dev = dict(DEVICE_SMALL, width=1e-5)
    seed, ns, dt = 8080, 20, 1e-15
    od = O.OracleDevice(dev)
    M = O.OracleModel(MODELS_DEVICE["device_para"])
    ref = od.run_philox(M, O.PM_CIC, 1e-14, 1.8, dt, ns, seed, O.thermal_voltage(300.0),
                        init_from_potential=True)
    ref_dop = od.run_philox(M, O.PM_CIC, 1e-14, 1.8, dt, 1, seed, O.thermal_voltage(300.0))
    print("potential init twin: %d particles (doping init: %d)"
          % (ref["n_init"][0], ref_dop["n_init"][0]), flush=True)
    assert ref["n_init"][0] != ref_dop["n_init"][0]
    ctx = emcgpu.Context(precision=emcgpu.FP64)
    ctx.setup(0, MODELS_DEVICE["device_para"])
    ctx.device_setup(dev, pm=emcgpu.PM_CIC, poisson_tol=1e-14, poisson_max_iter=80,
                     init_from_potential=True)
    assert ctx.device_init(seed) == ref["n_init"][0]
    out = ctx.device_step(dt, ns, seed=seed)
    assert (out["n_particles"] == ref["n_part"]).all()
    assert (out["exits"] == ref["exits"]).all()
    assert (out["injected"] - out["deleted"] == ref["injs"] - ref["dels"]).all()
    ctx.close()
