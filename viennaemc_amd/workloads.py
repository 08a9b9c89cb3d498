"""Model and workload descriptions shared by the golden harness, the oracle,
the GPU path, the tests and bench.py.

Each model mirrors one configuration built by ``oracle/ref_harness.cpp`` out of
the reference's own valley and mechanism classes.  The same dictionaries are
lowered to the oracle's handle API (``oracle/oracle_py.py``) and to the C-ABI
descriptors of the HIP library (``viennaemc_amd.emcgpu``), so every parity test
compares the three on identical physics.

Units: masses relative to m0, energies in eV, deformation potentials in eV
(acoustic) / eV per m (zero-order) / eV (first-order), temperatures in K,
fields in kV/cm.
"""

KB = 1.38066e-23  # include/emcConstants.hpp:15
Q = 1.60219e-19   # include/emcConstants.hpp:12
ME = 9.11e-31     # include/emcConstants.hpp:24

PARA_ISO, NONPARA_ISO, PARA_ANISO, NONPARA_ANISO = 0, 1, 2, 3
(ACOUSTIC, ZERO_IV_AB, ZERO_IV_EM, FIRST_IV_AB, FIRST_IV_EM, FROEHLICH_AB, FROEHLICH_EM,
 HOT_FROEHLICH_AB, HOT_FROEHLICH_EM, SCR_FROEHLICH_AB, SCR_FROEHLICH_EM, SCR_HOT_FROEHLICH_AB,
 SCR_HOT_FROEHLICH_EM) = range(13)

SI_RHO = 2329.0   # examples/SiliconFunctions.hpp:22
SI_VSOUND = 9040.0  # examples/SiliconFunctions.hpp:24

# X-valley frames of examples/SiliconFunctions.hpp:66-71
SI_AXES = [
    [(1, 0, 0), (0, 1, 0), (0, 0, 1)],
    [(0, 1, 0), (1, 0, 0), (0, 0, 1)],
    [(0, 0, 1), (0, 1, 0), (1, 0, 0)],
]
# [111]-type frames (rows need normalising; exercises setSubValleyEllipseCoordSystem)
AXES_111 = [
    [(1, 1, 1), (1, -1, 0), (1, 1, -2)],
    [(-1, 1, 1), (1, 1, 0), (-1, 1, -2)],
    [(1, -1, 1), (1, 1, 0), (1, -1, -2)],
    [(1, 1, -1), (1, -1, 0), (1, 1, 2)],
]
FINAL_G = [[0], [1], [2]]                            # SiliconFunctions.hpp:48-49
FINAL_F = [[1, 1, 2, 2], [0, 0, 2, 2], [0, 0, 1, 1]]  # SiliconFunctions.hpp:50-51

# Jacoboni & Reggiani, Rev. Mod. Phys. 55, 645 (1983): zero-order Si phonons
# (Theta [K], D [eV/cm]), g-type then f-type.
JAC_G = [(140, 0.5e8), (215, 0.8e8), (720, 11e8)]
JAC_F = [(220, 0.3e8), (550, 2e8), (685, 2e8)]


def theta_to_ev(theta):
    """kB * Theta / q, evaluated in the same order as the harness."""
    return KB * theta / Q


def _si_jacoboni_mechs():
    mechs = [dict(kind=ACOUSTIC, valley=0, sigma=9.0)]
    for theta, d in JAC_G:
        for kind in (ZERO_IV_AB, ZERO_IV_EM):
            mechs.append(dict(kind=kind, valley=0, final=0, defpot=d * 100.0,
                              hw=theta_to_ev(theta), fmap=FINAL_G))
    for theta, d in JAC_F:
        for kind in (ZERO_IV_AB, ZERO_IV_EM):
            mechs.append(dict(kind=kind, valley=0, final=0, defpot=d * 100.0,
                              hw=theta_to_ev(theta), fmap=FINAL_F))
    return mechs


def _si_vasileska_mechs():
    m = [dict(kind=ACOUSTIC, valley=0, sigma=9.0)]
    for kind in (ZERO_IV_AB, ZERO_IV_EM):
        m.append(dict(kind=kind, valley=0, final=0, defpot=5.23e10, hw=0.06, fmap=FINAL_F))
    for kind in (ZERO_IV_AB, ZERO_IV_EM):
        m.append(dict(kind=kind, valley=0, final=0, defpot=5.23e10, hw=0.06, fmap=FINAL_G))
    for kind in (FIRST_IV_AB, FIRST_IV_EM):
        m.append(dict(kind=kind, valley=0, final=0, defpot=2.5, hw=0.023, fmap=FINAL_F))
    for kind in (FIRST_IV_AB, FIRST_IV_EM):
        m.append(dict(kind=kind, valley=0, final=0, defpot=4.0, hw=0.018, fmap=FINAL_G))
    return m


def _two_valley_mechs():
    l_to_l = [[1, 2, 3], [0, 2, 3], [0, 1, 3], [0, 1, 2]]
    l_to_g = [[0], [0], [0], [0]]
    g_to_l = [[0, 1, 2, 3]]
    return [
        dict(kind=ACOUSTIC, valley=0, sigma=7.0),
        dict(kind=ZERO_IV_AB, valley=0, final=0, defpot=3e10, hw=0.029, fmap=l_to_l),
        dict(kind=ZERO_IV_EM, valley=0, final=0, defpot=3e10, hw=0.029, fmap=l_to_l),
        dict(kind=ZERO_IV_AB, valley=0, final=1, defpot=1e11, hw=0.0278, fmap=l_to_g),
        dict(kind=ZERO_IV_EM, valley=0, final=1, defpot=1e11, hw=0.0278, fmap=l_to_g),
        dict(kind=ACOUSTIC, valley=1, sigma=7.0),
        dict(kind=FIRST_IV_AB, valley=1, final=0, defpot=3.0, hw=0.0278, fmap=g_to_l),
        dict(kind=FIRST_IV_EM, valley=1, final=0, defpot=3.0, hw=0.0278, fmap=g_to_l),
    ]


MODELS = {
    # config 1/2 physics: Si X valleys, Jacoboni zero-order phonons, m_t = 0.19
    "si_jacoboni": dict(
        n_levels=2000, max_energy=2.0, T=300.0, field_kvcm=50.0, field_dir=(1, 0, 0),
        valleys=[dict(kind=NONPARA_ANISO, mass=(0.916, 0.19, 0.19), deg=3, alpha=0.5,
                      bottom=0.0, axes=SI_AXES)],
        mechs=_si_jacoboni_mechs()),
    # the reference's own bulk example (examples/bulkSimulation + SiliconFunctions.hpp)
    "si_vasileska": dict(
        n_levels=1000, max_energy=1.0, T=300.0, field_kvcm=10.0, field_dir=(-1, 0, 0),
        valleys=[dict(kind=NONPARA_ANISO, mass=(0.916, 0.196, 0.196), deg=3, alpha=0.5,
                      bottom=0.0, axes=SI_AXES)],
        mechs=_si_vasileska_mechs()),
    # config 5 physics (equilibrium phonons): one Kane Gamma valley, Froehlich only
    "polar_gamma": dict(
        n_levels=2000, max_energy=2.0, T=300.0, field_kvcm=20.0, field_dir=(0, 0, 1),
        valleys=[dict(kind=NONPARA_ISO, mass=(0.28, 0.28, 0.28), deg=1, alpha=0.3, bottom=0.0)],
        mechs=[dict(kind=FROEHLICH_AB, valley=0, hw=0.06, mrel=0.28, eps_hi=3.5, eps_lo=10.0),
               dict(kind=FROEHLICH_EM, valley=0, hw=0.06, mrel=0.28, eps_hi=3.5, eps_lo=10.0)]),
    # config 4 physics: one parabolic isotropic valley, m* = 0.26
    "para_iso": dict(
        n_levels=1000, max_energy=1.5, T=77.0, field_kvcm=5.0, field_dir=(1, 1, 0),
        valleys=[dict(kind=PARA_ISO, mass=(0.26, 0.26, 0.26), deg=1, alpha=0.0, bottom=0.0)],
        mechs=[dict(kind=ACOUSTIC, valley=0, sigma=9.0),
               dict(kind=ZERO_IV_AB, valley=0, final=0, defpot=5e10, hw=0.05, fmap=[[0]]),
               dict(kind=ZERO_IV_EM, valley=0, final=0, defpot=5e10, hw=0.05, fmap=[[0]])]),
    # two valleys: parabolic anisotropic [111] valley + raised Kane valley
    "two_valley": dict(
        n_levels=1500, max_energy=1.5, T=300.0, field_kvcm=30.0, field_dir=(1, 1, 1),
        valleys=[dict(kind=PARA_ANISO, mass=(1.6, 0.12, 0.12), deg=4, alpha=0.0, bottom=0.0,
                      axes=AXES_111),
                 dict(kind=NONPARA_ISO, mass=(0.067, 0.067, 0.067), deg=1, alpha=0.6,
                      bottom=0.3)],
        mechs=_two_valley_mechs()),
}

# hot-phonon model: one Kane Gamma valley whose Froehlich absorption and
# emission read and feed a 256-bin LO phonon bath (emcPhononBath), every
# particle starting at 0.5 eV (emcElectron's fixed-energy initialisation)
MODELS["hot_polar"] = dict(
    n_levels=2000, max_energy=2.0, T=300.0, field_kvcm=10.0, field_dir=(0, 0, 1), init_energy=0.5,
    valleys=[dict(kind=NONPARA_ISO, mass=(0.28, 0.28, 0.28), deg=1, alpha=0.3, bottom=0.0)],
    baths=[dict(n_bins=256, dq=2e7, tau_lo=1e-12, hw=0.06, volume=2.7e-23)],
    mechs=[dict(kind=HOT_FROEHLICH_AB, valley=0, hw=0.06, mrel=0.28, eps_hi=3.5, eps_lo=10.0, bath=0),
           dict(kind=HOT_FROEHLICH_EM, valley=0, hw=0.06, mrel=0.28, eps_hi=3.5, eps_lo=10.0, bath=0)])

# screened polar model: the Kane Gamma valley of polar_gamma with
# emcScreenedFroehlich{Absorption,Emission}3D sharing one
# emcPlasmonScreening at q_s^2 = 1e17 / m^2 (screening length 3.2 nm)
MODELS["screened_polar"] = dict(
    n_levels=2000, max_energy=2.0, T=300.0, field_kvcm=20.0, field_dir=(0, 0, 1),
    valleys=[dict(kind=NONPARA_ISO, mass=(0.28, 0.28, 0.28), deg=1, alpha=0.3, bottom=0.0)],
    screening=dict(qs2=1e17),
    mechs=[dict(kind=SCR_FROEHLICH_AB, valley=0, hw=0.06, mrel=0.28, eps_hi=3.5, eps_lo=10.0),
           dict(kind=SCR_FROEHLICH_EM, valley=0, hw=0.06, mrel=0.28, eps_hi=3.5, eps_lo=10.0)])

# screened hot-phonon model (the hotCarrierMHP default): the hot_polar bath
# read through emcScreenedHotPhononFroehlich{Absorption,Emission}3D with the
# q-resolved rate and angle; the screening (eps 3.5) starts at q_s^2 = 3e16
# / m^2 and in a run follows the carriers (n = N / 2.7e-23 m^3, T from the
# mean energy) after every step
MODELS["screened_hot_polar"] = dict(
    n_levels=2000, max_energy=2.0, T=300.0, field_kvcm=10.0, field_dir=(0, 0, 1), init_energy=0.5,
    valleys=[dict(kind=NONPARA_ISO, mass=(0.28, 0.28, 0.28), deg=1, alpha=0.3, bottom=0.0)],
    baths=[dict(n_bins=256, dq=2e7, tau_lo=1e-12, hw=0.06, volume=2.7e-23)],
    screening=dict(qs2=3e16, plasma=True, eps=3.5, volume=2.7e-23),
    mechs=[dict(kind=SCR_HOT_FROEHLICH_AB, valley=0, hw=0.06, mrel=0.28, eps_hi=3.5, eps_lo=10.0,
                bath=0, q_resolved=True, q_resolved_angle=True),
           dict(kind=SCR_HOT_FROEHLICH_EM, valley=0, hw=0.06, mrel=0.28, eps_hi=3.5, eps_lo=10.0,
                bath=0, q_resolved=True, q_resolved_angle=True)])

MODEL_NAMES = list(MODELS)

# the plain bulk Si run of the CPU reference path: the si_jacoboni physics
# (X valleys m_l 0.916 / m_t 0.19, alpha 0.5 /eV, acoustic + six zero-order
# Jacoboni phonons, 2000-level tables on (0, 2] eV) at 10 kV/cm along [100],
# 300 K, 10,000 electrons, 3000 steps of 1 fs, observables every step
# (examples/bulkSimulation/bulkSimulation.cpp:119-162)
C1_PARTICLES, C1_STEPS, C1_DT = 10_000, 3000, 1e-15
# steady-state window of the time averages (the last 2 ps) and the number of
# equal particle groups whose spread gives the statistical error
C1_AVG_FROM, C1_GROUPS = 1000, 20


def si_bulk_c1():
    import copy
    m = copy.deepcopy(MODELS["si_jacoboni"])
    m.update(field_kvcm=10.0, field_dir=(1, 0, 0))
    return m


def group_stats(per_group):
    """Mean over groups and its standard error, std(ddof=1) / sqrt(groups),
    along axis 0."""
    import numpy as np
    a = np.asarray(per_group, dtype=np.float64)
    return a.mean(axis=0), a.std(axis=0, ddof=1) / np.sqrt(a.shape[0])


def photoexcited_cooling(n_particles=4_000_000):
    """The hot-phonon cooling workload: 4M electrons in a Kane Gamma valley
    (m* 0.28, alpha 0.3/eV), energies drawn from N(0.5 eV, 0.05 eV), hot-
    phonon Froehlich absorption + emission (60 meV, eps 3.5 / 10) on a
    256-bin LO bath (dq 2e7 /m) whose LO phonons decay with tau_LO = 1 ps
    into the acoustic (Klemens) reservoir, hbar w_ac = hbar w_LO / 2, coupled
    to the lattice in 30 ps (the reference hot-carrier driver's defaults,
    examples/hotCarrierMHP/hotCarrierMHP.cpp:155, 209, 377), in the volume
    that holds them at 1e24 /m^3, no field, dt 0.5 fs."""
    import copy
    m = copy.deepcopy(MODELS["hot_polar"])
    m.update(field_kvcm=0.0, init_energy=0.5, init_sigma=0.05)
    b = m["baths"][0]
    b["volume"] = n_particles / 1e24
    b.update(acoustic=True, ac_hw=b["hw"] / 2, tau_ac=30e-12)
    return m


def screened_cooling(n_particles=4_000_000):
    """The cooling workload of photoexcited_cooling with the reference's
    hot-carrier default (examples/hotCarrierMHP/hotCarrierMHP.cpp:221, 415-435):
    screened hot-phonon Froehlich with the q-resolved rate and angle, and the
    plasma screening (eps 3.5) following the carriers after every step."""
    import copy
    m = photoexcited_cooling(n_particles)
    src = MODELS["screened_hot_polar"]
    m["mechs"] = copy.deepcopy(src["mechs"])
    m["screening"] = dict(src["screening"], volume=m["baths"][0]["volume"])
    return m


def field_force(model):
    """Force vector of the bulk handler: scale(normalize(dir) * E, -q)."""
    import math
    d = list(model["field_dir"])
    n = math.sqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2])
    d = [x / n for x in d]
    e = model["field_kvcm"] * 1e5
    return [x * e * -Q for x in d], d


def _splitmix64(state):
    """One SplitMix64 step: returns (new state, output)."""
    m = (1 << 64) - 1
    state = (state + 0x9E3779B97F4A7C15) & m
    z = state
    z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & m
    z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & m
    return state, z ^ (z >> 31)


def batched_instances(n_instances=64, seed=20261202, n_min=10_000, n_max=2_000_000):
    """The batched parameter sweep: independent Si bulk instances (the
    si_jacoboni physics) whose bias, temperature and size are drawn from a
    host SplitMix64 stream -- field log-uniform in [0.1, 200] kV/cm along
    [100] or [111] with probability 1/2 each, T uniform in [77, 500] K,
    particle count log-uniform in [n_min, n_max].  Each instance keeps its
    own FP64-built 2000-level tables.  Returns a list of (model, count)."""
    import copy
    import math
    state = seed
    out = []

    def uniform():
        nonlocal state
        state, z = _splitmix64(state)
        return (z >> 11) * 2.0 ** -53

    for _ in range(n_instances):
        field = math.exp(math.log(0.1) + (math.log(200.0) - math.log(0.1)) * uniform())
        direction = (1, 0, 0) if uniform() < 0.5 else (1, 1, 1)
        temperature = 77.0 + (500.0 - 77.0) * uniform()
        count = int(math.exp(math.log(n_min) + (math.log(n_max) - math.log(n_min)) * uniform()))
        m = copy.deepcopy(MODELS["si_jacoboni"])
        m.update(T=temperature, field_kvcm=field, field_dir=direction)
        out.append((m, count))
    return out


# ---- 2-D device (resistor2D-style) -------------------------------------------
SI_EPS_R = 11.8    # examples/SiliconFunctions.hpp:21
SI_NI = 1.45e16    # examples/SiliconFunctions.hpp:23

# one parabolic valley (m* 0.26) with acoustic + zero-order intervalley
# phonons at 300 K, registered for every doping region of the device
MODELS_DEVICE = {
    "device_para": dict(
        n_levels=1000, max_energy=1.5, T=300.0, field_kvcm=0.0, field_dir=(1, 0, 0),
        valleys=[dict(kind=PARA_ISO, mass=(0.26, 0.26, 0.26), deg=1, alpha=0.0, bottom=0.0)],
        mechs=[dict(kind=ACOUSTIC, valley=0, sigma=9.0),
               dict(kind=ZERO_IV_AB, valley=0, final=0, defpot=5e10, hw=0.05, fmap=[[0]]),
               dict(kind=ZERO_IV_EM, valley=0, final=0, defpot=5e10, hw=0.05, fmap=[[0]])]),
}


def device_region_specs(spec=None):
    """The device type with per-region scatter tables (addScatterMechanism
    with region lists): for doping regions 0 (the channel), 1 (source) and 2
    (drain) of DEVICE_SMALL the mechanisms registered there, in registration
    order: all three; acoustic only; acoustic and intervalley emission."""
    spec = spec or MODELS_DEVICE["device_para"]
    return [dict(spec, mechs=[spec["mechs"][i] for i in idx]) for idx in ([0, 1, 2], [0], [0, 2])]

# the 21 x 7 node n+ n n+ resistor of the golden fixtures: doping regions
# (min corner, max corner, doping [1/m^3]) in order, ohmic contacts
# (boundary 0=XMIN 1=XMAX 2=YMIN 3=YMAX, voltage, min, max along the boundary)
DEVICE_SMALL = dict(
    max_pos=(20e-9, 6e-9), spacing=(1e-9, 1e-9), width=1e-6, T=300.0, eps_r=SI_EPS_R, ni=SI_NI,
    doping=[((0, 0), (20e-9, 6e-9), 1e24), ((0, 0), (3e-9, 6e-9), 1e25),
            ((17e-9, 0), (20e-9, 6e-9), 1e25)],
    contacts=[(0, 0.2, 0.0, 6e-9), (1, 0.0, 0.0, 6e-9)])


# the same resistor with a gate (Robin) contact on each y boundary: (boundary,
# voltage, min, max, relative permittivity of the oxide, oxide thickness [m],
# barrier height [V]) as in addGateContact (emcDevice.hpp:195-202)
DEVICE_GATED = dict(DEVICE_SMALL, contacts=DEVICE_SMALL["contacts"] + [
    (3, 0.4, 5e-9, 15e-9, 3.9, 1e-9, 0.3),
    (2, -0.2, 8e-9, 12e-9, 25.0, 2e-9, 0.1)])

SI_BANDGAP = 1.15  # examples/SiliconFunctions.hpp:25 [eV]


def mosfet_device(vg=1.0, vd=1.0, width=1e-6):
    """The n-channel MOSFET of the mosfet2D example (mosfet2D.cpp:62-87):
    125 x 100 nm at 1 nm (126 x 101 nodes), p-type bulk 5e23 /m^3 below
    30 nm, n+ source [0, 51] nm and drain [75, 125] nm at 5e25 /m^3, p-type
    5e24 /m^3 under the gate; ohmic substrate (YMAX), source and drain (YMIN,
    drain at vd), a gate on YMIN over [51, 75] nm (oxide eps_r 3.9, 1.2 nm,
    barrier half the band gap) at vg.  The reference runs it at a width of
    1 um; a smaller width scales the particle count down."""
    return dict(max_pos=(125e-9, 100e-9), spacing=(1e-9, 1e-9), width=width, T=300.0,
                eps_r=SI_EPS_R, ni=SI_NI,
                doping=[((0, 30e-9), (125e-9, 100e-9), -5e23), ((0, 0), (51e-9, 30e-9), 5e25),
                        ((51e-9, 0), (75e-9, 30e-9), -5e24), ((75e-9, 0), (125e-9, 30e-9), 5e25)],
                contacts=[(3, 0.0, 0.0, 125e-9), (2, 0.0, 0.0, 51e-9),
                          (2, vg, 51e-9, 75e-9, 3.9, 1.2e-9, SI_BANDGAP / 2),
                          (2, vd, 75e-9, 125e-9)])


def resistor_device(nx=200, ny=50, h=1e-9, n_particles=2_000_000, contact_nodes=10, bias=0.5):
    """The 2-D resistor workload: nx x ny nodes at spacing h, N_D = 1e23 /m^3
    with 1e25 /m^3 ohmic contact regions of contact_nodes columns at both
    ends, bias on the XMIN contact, and the device width that makes the
    doping-weighted particle count about n_particles."""
    mx, my = (nx - 1) * h, (ny - 1) * h
    xc = (contact_nodes - 1) * h
    dev = dict(max_pos=(mx, my), spacing=(h, h), width=1e-6, T=300.0, eps_r=SI_EPS_R, ni=SI_NI,
               doping=[((0, 0), (mx, my), 1e23), ((0, 0), (xc, my), 1e25),
                       ((mx - xc, 0), (mx, my), 1e25)],
               contacts=[(0, bias, 0.0, my), (1, 0.0, 0.0, my)])
    # sum of doping * (1/2 per boundary dimension) over the nodes, per m of width
    w = 0.0
    for iy in range(ny):
        for ix in range(nx):
            d = 1e25 if (ix < contact_nodes or ix >= nx - contact_nodes) else 1e23
            for c, n in ((ix, nx), (iy, ny)):
                if c == 0 or c == n - 1:
                    d *= 0.5
            w += d * h * h
    dev["width"] = n_particles / w
    return dev


def run_groups_gpu(emcgpu, spec, precision, groups, per_group, steps, dt, seed, avg_from,
                   steps_per_launch=0, ppl=0):
    """One GPU context holding `groups` instances of `per_group` particles of
    the same model (each instance its own Philox streams): thermal
    initialisation, `steps` steps, and per group the time averages over steps
    [avg_from, steps) of the valley-0 drift velocity, the valley-0 mean
    energy and the occupancy of every subvalley of valley 0.  Returns
    dict(v=[groups], E=[groups], sub=[groups, deg], wall=seconds of the
    step call, ctx_steps_per_launch)."""
    import time
    import numpy as np
    ctx = emcgpu.Context(precision=precision, n_instances=groups)
    for g in range(groups):
        ctx.setup(g, spec)
    ctx.resize([per_group] * groups)
    for g in range(groups):
        ctx.init_particles(g, seed, spec["T"])
    ctx.synchronize()
    t0 = time.perf_counter()
    ctx.step(dt, steps, seed=seed, steps_per_launch=steps_per_launch, ppl=ppl)
    ctx.synchronize()
    wall = time.perf_counter() - t0
    deg = spec["valleys"][0]["deg"]
    v, E, sub = [], [], []
    for g in range(groups):
        av = ctx.averages(g)
        v.append(av["avgV"][avg_from:, 0].mean())
        E.append(av["avgE"][avg_from:, 0].mean())
        sub.append(av["sub_occ"][avg_from:, 0, :deg].mean(axis=0))
    ctx.close()
    return dict(v=np.array(v), E=np.array(E), sub=np.array(sub), wall=wall)
