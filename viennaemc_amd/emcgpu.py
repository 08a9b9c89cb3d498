"""ctypes binding of the C ABI in include/emcgpu.h (libemcgpu.so).

The HIP library is the only compute path: importing this module does not
need a GPU, but creating a :class:`Context` loads ``libemcgpu.so`` and fails
loudly if it is missing or if no HIP device is present.  There is no CPU
fallback.
"""
import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libemcgpu.so")

MAX_VALLEYS, MAX_DEG, MAX_Z, MAX_MECH, MAX_LEVELS = 4, 8, 8, 32, 8192
MAX_BATHS, MAX_BINS = 4, 1024
FP64, FP32 = 0, 1
PRODUCTION, REPLAY = 0, 1
PARA_ISO, NONPARA_ISO, PARA_ANISO, NONPARA_ANISO = 0, 1, 2, 3
(ACOUSTIC, ZERO_IV_AB, ZERO_IV_EM, FIRST_IV_AB, FIRST_IV_EM, FROEHLICH_AB,
 FROEHLICH_EM, HOT_FROEHLICH_AB, HOT_FROEHLICH_EM, SCREENED_FROEHLICH_AB,
 SCREENED_FROEHLICH_EM, SCREENED_HOT_FROEHLICH_AB, SCREENED_HOT_FROEHLICH_EM) = range(13)
SCREENING_FIXED, SCREENING_PLASMA = 0, 1

PM_NGP, PM_CIC, PM_NEC = 0, 1, 2
XMIN, XMAX, YMIN, YMAX = 0, 1, 2, 3
MAX_CONTACTS, MAX_REGIONS = 8, 16

STATUS = {0: "OK", 1: "ERR_ARG", 2: "ERR_HIP", 3: "ERR_STATE", 4: "ERR_DRAWS", 5: "ERR_NOMEM"}

# include/emcConstants.hpp:9-25
Q = 1.60219e-19
ME = 9.11e-31
KB = 1.38066e-23


class EmcGpuError(RuntimeError):
    def __init__(self, status, msg):
        super().__init__("%s: %s" % (STATUS.get(status, status), msg))
        self.status = status


class Config(C.Structure):
    _fields_ = [("device", C.c_int), ("precision", C.c_int), ("n_instances", C.c_int)]


class ValleyDesc(C.Structure):
    _fields_ = [("kind", C.c_int), ("rel_mass", C.c_double * 3), ("particle_mass", C.c_double),
                ("degeneracy", C.c_int), ("alpha", C.c_double), ("bottom_energy", C.c_double),
                ("axes", (C.c_double * 9) * MAX_DEG)]


class MechDesc(C.Structure):
    _fields_ = [("kind", C.c_int), ("valley", C.c_int), ("final_valley", C.c_int),
                ("phonon_energy", C.c_double), ("coupling", C.c_double),
                ("rel_eff_mass", C.c_double), ("eps_hi", C.c_double), ("eps_lo", C.c_double),
                ("z", C.c_int), ("final_sub", (C.c_int * MAX_Z) * MAX_DEG), ("bath", C.c_int),
                ("q_resolved", C.c_int), ("q_resolved_angle", C.c_int)]


class BathDesc(C.Structure):
    _fields_ = [("n_bins", C.c_int), ("dq", C.c_double), ("tau_lo", C.c_double),
                ("phonon_energy", C.c_double), ("lattice_temp", C.c_double), ("volume", C.c_double),
                ("acoustic", C.c_int), ("ac_phonon_energy", C.c_double), ("tau_acoustic", C.c_double),
                ("w_ridley", C.c_double), ("to_phonon_energy", C.c_double), ("tau_to", C.c_double),
                ("tau_profile", C.c_void_p)]


class ScreeningDesc(C.Structure):
    _fields_ = [("mode", C.c_int), ("enabled", C.c_int), ("eps_static", C.c_double),
                ("qs2", C.c_double), ("volume", C.c_double)]


class Material(C.Structure):
    _fields_ = [("eps_r", C.c_double), ("rho", C.c_double), ("ni", C.c_double),
                ("v_sound", C.c_double)]


class StepArgs(C.Structure):
    _fields_ = [("dt", C.c_double), ("n_steps", C.c_int), ("steps_per_launch", C.c_int),
                ("mode", C.c_int), ("seed", C.c_uint64), ("step0", C.c_uint32),
                ("timing", C.c_int), ("particles_per_lane", C.c_int), ("use_graph", C.c_int),
                ("bath_record_only", C.c_int), ("rebuild_every", C.c_int), ("draws", C.c_void_p),
                ("draw_stride", C.c_size_t),
                ("cursor", C.c_void_p), ("grain_tau", C.c_void_p),
                ("trace_valley", C.c_void_p), ("trace_sub", C.c_void_p),
                ("trace_events", C.c_void_p), ("trace_k", C.c_void_p),
                ("trace_energy", C.c_void_p), ("trace_pos", C.c_void_p)]


class DopingRegion(C.Structure):
    _fields_ = [("min_pos", C.c_double * 2), ("max_pos", C.c_double * 2), ("doping", C.c_double)]


CONTACT_OHMIC, CONTACT_SCHOTTKY, CONTACT_GATE = 0, 1, 2


class ContactDesc(C.Structure):
    _fields_ = [("boundary", C.c_int), ("voltage", C.c_double), ("min_pos", C.c_double),
                ("max_pos", C.c_double), ("type", C.c_int), ("eps_oxide", C.c_double),
                ("thickness", C.c_double), ("barrier_height", C.c_double)]


class DeviceDesc(C.Structure):
    _fields_ = [("max_pos", C.c_double * 2), ("spacing", C.c_double * 2), ("width", C.c_double),
                ("temperature", C.c_double), ("eps_r", C.c_double), ("ni", C.c_double),
                ("n_doping", C.c_int), ("doping", C.c_void_p), ("n_contacts", C.c_int),
                ("contacts", C.c_void_p), ("pm_scheme", C.c_int), ("poisson_tol", C.c_double),
                ("poisson_max_iter", C.c_int), ("init_from_potential", C.c_int),
                ("region_tables", C.c_int)]


class Grid3(C.Structure):
    """emcgpu_grid3: a 3-D grid of the particle-mesh operators."""
    _fields_ = [("ext", C.c_int * 3), ("spacing", C.c_double * 3), ("pm_scheme", C.c_int)]


class DeviceDriftArgs(C.Structure):
    _fields_ = [("dt", C.c_double), ("n_steps", C.c_int), ("mode", C.c_int),
                ("seed", C.c_uint64), ("step0", C.c_uint32), ("draws", C.c_void_p),
                ("draw_stride", C.c_longlong), ("cursor", C.c_void_p), ("grain_tau", C.c_void_p),
                ("exit_step", C.c_void_p), ("exit_contact", C.c_void_p), ("tr_pos", C.c_void_p),
                ("tr_k", C.c_void_p), ("tr_E", C.c_void_p), ("tr_events", C.c_void_p)]


class DeviceStepArgs(C.Structure):
    _fields_ = [("dt", C.c_double), ("n_steps", C.c_int), ("seed", C.c_uint64),
                ("use_graph", C.c_int), ("exits", C.c_void_p), ("deleted", C.c_void_p),
                ("injected", C.c_void_p), ("n_particles", C.c_void_p),
                ("poisson_iters", C.c_void_p), ("poisson_resid", C.c_void_p),
                ("checksum", C.c_void_p), ("poisson_interval", C.c_int),
                ("channel_x0", C.c_double), ("channel_x1", C.c_double),
                ("channel_length", C.c_double), ("channel_current", C.c_void_p),
                ("accumulate", C.c_int), ("poisson_mode", C.c_int)]


POISSON_AUTO, POISSON_MULTIGRID, POISSON_DENSE = 0, 1, 2


TUNE_MAX = 32


class AutotuneArgs(C.Structure):
    _fields_ = [("dt", C.c_double), ("n_steps", C.c_int), ("seed", C.c_uint64),
                ("step0", C.c_uint32), ("repeats", C.c_int), ("n_ppl", C.c_int),
                ("ppl", C.c_int * 4), ("n_spl", C.c_int), ("spl", C.c_int * 8),
                ("n_results", C.c_int), ("res_ppl", C.c_int * TUNE_MAX),
                ("res_spl", C.c_int * TUNE_MAX), ("res_ms_per_step", C.c_double * TUNE_MAX),
                ("best_ppl", C.c_int), ("best_spl", C.c_int), ("best_ms_per_step", C.c_double)]


_LIB = None
EXPORTS = [
    "emcgpu_create", "emcgpu_destroy", "emcgpu_last_error", "emcgpu_set_model",
    "emcgpu_build_tables", "emcgpu_set_rate_tables", "emcgpu_get_tables", "emcgpu_set_field",
    "emcgpu_resize", "emcgpu_get_count", "emcgpu_set_particles", "emcgpu_get_particles",
    "emcgpu_init_particles", "emcgpu_step", "emcgpu_get_observables", "emcgpu_obs_layout",
    "emcgpu_get_timing", "emcgpu_synchronize", "emcgpu_init_particles_energy", "emcgpu_set_baths",
    "emcgpu_get_bath", "emcgpu_set_bath_state", "emcgpu_get_bath_tallies", "emcgpu_get_bath_history",
    "emcgpu_device_setup", "emcgpu_device_info", "emcgpu_device_grids", "emcgpu_poisson",
    "emcgpu_efield", "emcgpu_deposit", "emcgpu_interpolate", "emcgpu_contacts",
    "emcgpu_device_set_particles", "emcgpu_device_get_particles", "emcgpu_device_set_efield",
    "emcgpu_device_drift", "emcgpu_device_init", "emcgpu_device_step", "emcgpu_device_fields",
    "emcgpu_autotune", "emcgpu_device_averages", "emcgpu_set_screening", "emcgpu_get_screening",
    "emcgpu_update_screening", "emcgpu_build_tables_all", "emcgpu_step_noalloc",
    "emcgpu_reserve", "emcgpu_get_bath_energy", "emcgpu_set_box", "emcgpu_set_positions",
    "emcgpu_get_positions", "emcgpu_observe", "emcgpu_deposit3", "emcgpu_interpolate3",
    "emcgpu_efield3",
]


def load(path=LIB_PATH):
    """Load libemcgpu.so and declare the C signatures (no device calls)."""
    global _LIB
    if _LIB is not None:
        return _LIB
    if not os.path.exists(path):
        raise RuntimeError("libemcgpu.so is not built (%s); run __graft_entry__.build()" % path)
    L = C.CDLL(path)
    vp, i, d, p = C.c_void_p, C.c_int, C.c_double, C.POINTER
    sig = {
        "emcgpu_create": (i, [p(Config), p(vp)]),
        "emcgpu_destroy": (None, [vp]),
        "emcgpu_last_error": (C.c_char_p, [vp]),
        "emcgpu_set_model": (i, [vp, i, p(ValleyDesc), i, p(MechDesc), i]),
        "emcgpu_build_tables": (i, [vp, i, i, d, d, p(Material)]),
        "emcgpu_set_rate_tables": (i, [vp, i, i, d, vp]),
        "emcgpu_build_tables_all": (i, [vp, i, d, p(d), vp]),
        "emcgpu_get_tables": (i, [vp, i, i, vp, p(d), p(i)]),
        "emcgpu_set_field": (i, [vp, i, p(d), d, d]),
        "emcgpu_resize": (i, [vp, p(C.c_size_t)]),
        "emcgpu_get_count": (i, [vp, i, p(C.c_size_t)]),
        "emcgpu_set_particles": (i, [vp, i, vp, vp, vp, vp, vp]),
        "emcgpu_get_particles": (i, [vp, i, vp, vp, vp, vp, vp]),
        "emcgpu_init_particles": (i, [vp, i, C.c_uint64, d]),
        "emcgpu_step": (i, [vp, p(StepArgs)]),
        "emcgpu_step_noalloc": (i, [vp, p(StepArgs), vp, C.c_size_t]),
        "emcgpu_reserve": (i, [vp, p(StepArgs)]),
        "emcgpu_set_box": (i, [vp, i, p(d)]),
        "emcgpu_set_positions": (i, [vp, i, vp]),
        "emcgpu_get_positions": (i, [vp, i, vp]),
        "emcgpu_observe": (i, [vp, vp]),
        "emcgpu_get_observables": (i, [vp, i, vp, p(i), p(i)]),
        "emcgpu_obs_layout": (i, [vp, p(i), p(i), p(i)]),
        "emcgpu_get_timing": (i, [vp, p(d), p(C.c_longlong), i]),
        "emcgpu_synchronize": (i, [vp]),
        "emcgpu_init_particles_energy": (i, [vp, i, C.c_uint64, d, d]),
        "emcgpu_set_baths": (i, [vp, i, p(BathDesc), i]),
        "emcgpu_set_screening": (i, [vp, i, p(ScreeningDesc)]),
        "emcgpu_get_screening": (i, [vp, i, p(C.c_double)]),
        "emcgpu_update_screening": (i, [vp, i]),
        "emcgpu_get_bath": (i, [vp, i, i, vp, vp]),
        "emcgpu_set_bath_state": (i, [vp, i, i, vp, d, d]),
        "emcgpu_get_bath_tallies": (i, [vp, i, i, vp, vp]),
        "emcgpu_get_bath_history": (i, [vp, i, i, vp, p(i)]),
        "emcgpu_get_bath_energy": (i, [vp, i, i, vp, p(i)]),
        "emcgpu_device_setup": (i, [vp, p(DeviceDesc)]),
        "emcgpu_device_info": (i, [vp, vp, vp]),
        "emcgpu_device_grids": (i, [vp, vp, vp, vp]),
        "emcgpu_poisson": (i, [vp, vp, vp, i, i, p(i), p(d)]),
        "emcgpu_efield": (i, [vp, vp, vp, vp]),
        "emcgpu_deposit": (i, [vp, vp, C.c_size_t, vp]),
        "emcgpu_interpolate": (i, [vp, vp, vp, vp, C.c_size_t, d, vp]),
        "emcgpu_deposit3": (i, [vp, p(Grid3), vp, C.c_size_t, vp]),
        "emcgpu_interpolate3": (i, [vp, p(Grid3), vp, vp, vp, vp, C.c_size_t, d, vp]),
        "emcgpu_efield3": (i, [vp, p(Grid3), d, vp, vp, vp, vp, vp]),
        "emcgpu_contacts": (i, [vp, vp, C.c_size_t, vp, vp]),
        "emcgpu_device_set_particles": (i, [vp, C.c_size_t, vp, vp, vp, vp, vp, vp]),
        "emcgpu_device_get_particles": (i, [vp, p(C.c_size_t), vp, vp, vp, vp, vp, vp]),
        "emcgpu_device_set_efield": (i, [vp, vp, vp]),
        "emcgpu_device_drift": (i, [vp, p(DeviceDriftArgs)]),
        "emcgpu_device_init": (i, [vp, C.c_uint64, p(C.c_size_t)]),
        "emcgpu_device_step": (i, [vp, p(DeviceStepArgs)]),
        "emcgpu_device_fields": (i, [vp, vp, vp, vp, vp, vp]),
        "emcgpu_autotune": (i, [vp, p(AutotuneArgs)]),
        "emcgpu_device_averages": (i, [vp, vp, vp, p(C.c_longlong), i]),
    }
    for name, (res, args) in sig.items():
        f = getattr(L, name)
        f.restype = res
        f.argtypes = args
    _LIB = L
    return L


def _ptr(a):
    return None if a is None else a.ctypes.data_as(C.c_void_p)


def _need_size(name, a, n):
    """The C ABI takes host buffers by pointer: their sizes are checked here."""
    if a.size != n:
        raise ValueError("%s holds %d values, the grid needs %d" % (name, a.size, n))


def valley_desc(v, particle_mass=ME):
    """Lower a valley dictionary (viennaemc_amd/workloads.py format) to the C descriptor."""
    d = ValleyDesc()
    d.kind = v["kind"]
    for j in range(3):
        d.rel_mass[j] = v["mass"][j]
    d.particle_mass = particle_mass
    d.degeneracy = v["deg"]
    d.alpha = v.get("alpha", 0.0)
    d.bottom_energy = v.get("bottom", 0.0)
    for s, ax in enumerate(v.get("axes", [])):
        flat = [x for row in ax for x in row]
        for j in range(9):
            d.axes[s][j] = flat[j]
    return d


def mech_desc(m):
    """Lower a mechanism dictionary (viennaemc_amd/workloads.py format) to the C descriptor."""
    d = MechDesc()
    d.kind = m["kind"]
    d.valley = m["valley"]
    d.final_valley = m.get("final", m["valley"])
    if m["kind"] == ACOUSTIC:
        d.coupling = m["sigma"]
    elif m["kind"] in (ZERO_IV_AB, ZERO_IV_EM, FIRST_IV_AB, FIRST_IV_EM):
        d.coupling = m["defpot"]
        d.phonon_energy = m["hw"]
        fmap = m["fmap"]
        d.z = len(fmap[0])
        for s, row in enumerate(fmap):
            for z, f in enumerate(row):
                d.final_sub[s][z] = f
    else:
        d.phonon_energy = m["hw"]
        d.rel_eff_mass = m["mrel"]
        d.eps_hi = m["eps_hi"]
        d.eps_lo = m["eps_lo"]
        d.bath = m.get("bath", 0)
        d.q_resolved = int(bool(m.get("q_resolved", False)))
        d.q_resolved_angle = int(bool(m.get("q_resolved_angle", True)))
    return d


def bath_desc(b, T):
    """Lower a bath dictionary (viennaemc_amd/workloads.py format) to the C descriptor;
    the returned profile array must outlive the call that uses the descriptor."""
    d = BathDesc()
    d.n_bins = b["n_bins"]
    d.dq = b["dq"]
    d.tau_lo = b["tau_lo"]
    d.phonon_energy = b["hw"]
    d.lattice_temp = b.get("T", T)
    d.volume = b["volume"]
    d.acoustic = int(b.get("acoustic", False))
    d.ac_phonon_energy = b.get("ac_hw", 0.0)
    d.tau_acoustic = b.get("tau_ac", 0.0)
    d.w_ridley = b.get("w_ridley", 0.0)
    d.to_phonon_energy = b.get("to_hw", 0.0)
    d.tau_to = b.get("tau_to", 0.0)
    prof = None
    if b.get("tau_profile") is not None:
        prof = np.ascontiguousarray(b["tau_profile"], dtype=np.float64)
        d.tau_profile = _ptr(prof)
    return d, prof


class Context:
    """One emcgpu context: particle state + tables on one GPU."""

    def __init__(self, precision=FP64, n_instances=1, device=0):
        self.L = load()
        cfg = Config(device, precision, n_instances)
        h = C.c_void_p()
        st = self.L.emcgpu_create(C.byref(cfg), C.byref(h))
        if st:
            raise EmcGpuError(st, self.L.emcgpu_last_error(None).decode())
        self.h = h
        self.precision = precision
        self.n_instances = n_instances
        self.counts = [0] * n_instances
        self.n_valleys = [0] * n_instances
        self.tracks_positions = False

    def close(self):
        if getattr(self, "h", None):
            self.L.emcgpu_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _chk(self, st):
        if st:
            raise EmcGpuError(st, self.L.emcgpu_last_error(self.h).decode())

    # -- model -----------------------------------------------------------
    def set_model(self, instance, spec, particle_mass=ME):
        vs = spec["valleys"]
        ms = spec["mechs"]
        va = (ValleyDesc * len(vs))(*[valley_desc(v, particle_mass) for v in vs])
        ma = (MechDesc * max(len(ms), 1))(*[mech_desc(m) for m in ms])
        self._chk(self.L.emcgpu_set_model(self.h, instance, va, len(vs), ma, len(ms)))
        self.n_valleys[instance] = len(vs)

    def build_tables(self, instance, n_levels, max_energy, T, rho=2329.0, v_sound=9040.0,
                     eps_r=11.8, ni=1.45e16):
        mat = Material(eps_r, rho, ni, v_sound)
        self._chk(self.L.emcgpu_build_tables(self.h, instance, n_levels, max_energy, T,
                                             C.byref(mat)))

    def build_tables_all(self, n_levels, max_energy, T, materials=None):
        """Tables of every instance with the rates evaluated on the GPU.
        T: one temperature or one per instance; materials: None (silicon's
        constants) or a list of dicts with rho / v_sound / eps_r / ni."""
        n = self.n_instances
        temps = np.ascontiguousarray(np.broadcast_to(np.asarray(T, dtype=np.float64), (n,)))
        mats = (Material * n)()
        for k in range(n):
            m = (materials or [{}] * n)[k]
            mats[k] = Material(m.get("eps_r", 11.8), m.get("rho", 2329.0), m.get("ni", 1.45e16),
                               m.get("v_sound", 9040.0))
        self._chk(self.L.emcgpu_build_tables_all(self.h, n_levels, max_energy,
                                                 temps.ctypes.data_as(C.POINTER(C.c_double)),
                                                 C.cast(mats, C.c_void_p)))

    def set_rate_tables(self, instance, max_energy, rates):
        r = np.ascontiguousarray(rates, dtype=np.float64)
        self._chk(self.L.emcgpu_set_rate_tables(self.h, instance, r.shape[1], max_energy, _ptr(r)))

    def get_tables(self, instance, valley, n_levels):
        nt = C.c_int()
        tau0 = C.c_double()
        self._chk(self.L.emcgpu_get_tables(self.h, instance, valley, None, C.byref(tau0),
                                           C.byref(nt)))
        out = np.zeros(nt.value * n_levels)
        self._chk(self.L.emcgpu_get_tables(self.h, instance, valley, _ptr(out), C.byref(tau0),
                                           C.byref(nt)))
        return out.reshape(nt.value, n_levels), tau0.value

    def set_field(self, instance, direction, strength, charge=-Q):
        d = (C.c_double * 3)(*direction)
        self._chk(self.L.emcgpu_set_field(self.h, instance, d, strength, charge))

    def set_baths(self, instance, baths, T):
        lowered = [bath_desc(b, T) for b in baths]
        arr = (BathDesc * max(len(lowered), 1))(*[d for d, _ in lowered])
        self._chk(self.L.emcgpu_set_baths(self.h, instance, arr, len(lowered)))
        self.n_bins = getattr(self, "n_bins", {})
        for b, bd in enumerate(baths):
            self.n_bins[(instance, b)] = bd["n_bins"]

    def get_bath(self, instance=0, bath=0):
        nb = self.n_bins[(instance, bath)]
        nq, sc = np.zeros(nb), np.zeros(6)
        self._chk(self.L.emcgpu_get_bath(self.h, instance, bath, _ptr(nq), _ptr(sc)))
        return dict(nq=nq, mean=sc[0], n_ac=sc[1], n_to=sc[2], n0=sc[3], t_ac=sc[4], t_to=sc[5])

    def set_bath_state(self, instance, bath, nq, n_ac=0.0, n_to=0.0):
        nq = np.ascontiguousarray(nq, dtype=np.float64)
        self._chk(self.L.emcgpu_set_bath_state(self.h, instance, bath, _ptr(nq), n_ac, n_to))

    def bath_tallies(self, instance=0, bath=0):
        nb = self.n_bins[(instance, bath)]
        em, ab = np.zeros(nb, dtype=np.uint64), np.zeros(nb, dtype=np.uint64)
        self._chk(self.L.emcgpu_get_bath_tallies(self.h, instance, bath, _ptr(em), _ptr(ab)))
        return em, ab

    def bath_history(self, instance=0, bath=0):
        n = C.c_int()
        self._chk(self.L.emcgpu_get_bath_history(self.h, instance, bath, None, C.byref(n)))
        out = np.zeros(n.value)
        self._chk(self.L.emcgpu_get_bath_history(self.h, instance, bath, _ptr(out), C.byref(n)))
        return out

    def bath_energy(self, instance=0, bath=0):
        """Energy record of every bath update of the last step() call, shape
        (n_steps, 3) in eV: net emitted phonon energy, LO energy gained, LO
        energy passed to the reservoirs."""
        n = C.c_int()
        self._chk(self.L.emcgpu_get_bath_energy(self.h, instance, bath, None, C.byref(n)))
        out = np.zeros(3 * n.value)
        self._chk(self.L.emcgpu_get_bath_energy(self.h, instance, bath, _ptr(out), C.byref(n)))
        return out.reshape(n.value, 3)

    def set_screening(self, instance, qs2, enabled=True, eps_static=1.0, mode=SCREENING_FIXED,
                      volume=0.0):
        d = ScreeningDesc(mode, 1 if enabled else 0, eps_static, qs2, volume)
        self._chk(self.L.emcgpu_set_screening(self.h, instance, C.byref(d)))

    def update_screening(self, instance=0):
        """Plasma screening: q_s^2 from the current particles, tables rebuilt."""
        self._chk(self.L.emcgpu_update_screening(self.h, instance))

    def get_screening(self, instance=0):
        v = C.c_double()
        self._chk(self.L.emcgpu_get_screening(self.h, instance, C.byref(v)))
        return v.value

    def setup(self, instance, spec, T=None):
        """set_model (+ set_baths, set_screening) + build_tables + set_field from a
        model dictionary."""
        T = spec["T"] if T is None else T
        self.set_model(instance, spec)
        if spec.get("baths"):
            self.set_baths(instance, spec["baths"], T)
        scr = spec.get("screening")
        if scr:
            if scr.get("plasma"):
                self.set_screening(instance, scr.get("qs2", 0.0), eps_static=scr["eps"],
                                   mode=SCREENING_PLASMA, volume=scr["volume"])
            else:
                self.set_screening(instance, scr["qs2"])
        self.build_tables(instance, spec["n_levels"], spec["max_energy"], T,
                          rho=spec.get("rho", 2329.0), v_sound=spec.get("v_sound", 9040.0))
        self.set_field(instance, spec["field_dir"], spec["field_kvcm"] * 1e5)

    # -- particles -------------------------------------------------------
    def resize(self, counts):
        arr = (C.c_size_t * self.n_instances)(*counts)
        self._chk(self.L.emcgpu_resize(self.h, arr))
        self.counts = list(counts)

    def set_particles(self, instance, k, E, tau, valley, sub):
        k = np.ascontiguousarray(k, dtype=np.float64).reshape(-1)
        E = np.ascontiguousarray(E, dtype=np.float64)
        tau = np.ascontiguousarray(tau, dtype=np.float64)
        valley = np.ascontiguousarray(valley, dtype=np.int32)
        sub = np.ascontiguousarray(sub, dtype=np.int32)
        n = self.counts[instance]
        if k.size != 3 * n or E.size != n or tau.size != n or valley.size != n or sub.size != n:
            raise ValueError("particle arrays must hold %d particles" % n)
        self._chk(self.L.emcgpu_set_particles(self.h, instance, _ptr(k), _ptr(E), _ptr(tau),
                                              _ptr(valley), _ptr(sub)))

    def get_particles(self, instance=0):
        n = self.counts[instance]
        k = np.zeros(3 * n)
        E = np.zeros(n)
        tau = np.zeros(n)
        v = np.zeros(n, dtype=np.int32)
        s = np.zeros(n, dtype=np.int32)
        self._chk(self.L.emcgpu_get_particles(self.h, instance, _ptr(k), _ptr(E), _ptr(tau),
                                              _ptr(v), _ptr(s)))
        return dict(k=k, E=E, tau=tau, valley=v, sub=s)

    def init_particles(self, instance, seed, T):
        self._chk(self.L.emcgpu_init_particles(self.h, instance, seed, T))

    def init_particles_energy(self, instance, seed, mean, sigma=0.0):
        self._chk(self.L.emcgpu_init_particles_energy(self.h, instance, seed, mean, sigma))

    # -- stepping --------------------------------------------------------
    def step(self, dt, n_steps, mode=PRODUCTION, seed=0, step0=0, steps_per_launch=0,
             timing=False, draws=None, cursor=None, grain_tau=None, trace=False, ppl=0,
             graph=False, bath_record_only=False, rebuild_every=0):
        a = StepArgs()
        a.dt = dt
        a.n_steps = n_steps
        a.steps_per_launch = steps_per_launch
        a.mode = mode
        a.seed = seed
        a.step0 = step0
        a.timing = 1 if timing else 0
        a.particles_per_lane = ppl
        a.use_graph = 1 if graph else 0
        a.bath_record_only = 1 if bath_record_only else 0
        a.rebuild_every = rebuild_every
        keep = []
        if mode == REPLAY:
            draws = np.ascontiguousarray(draws, dtype=np.uint64)
            if cursor.dtype != np.uint64 or not cursor.flags["C_CONTIGUOUS"]:
                raise ValueError("cursor must be a contiguous uint64 array")
            a.draws = _ptr(draws)
            a.draw_stride = draws.shape[1]
            a.cursor = _ptr(cursor)
            if grain_tau is not None:
                a.grain_tau = _ptr(grain_tau)
            keep += [draws]
        tr = None
        if trace:
            n = self.counts[0]
            tr = dict(valley=np.zeros((n_steps, n), dtype=np.int32),
                      sub=np.zeros((n_steps, n), dtype=np.int32),
                      events=np.zeros((n_steps, n), dtype=np.int32),
                      k=np.zeros((n_steps, n, 3)), E=np.zeros((n_steps, n)))
            a.trace_valley = _ptr(tr["valley"])
            a.trace_sub = _ptr(tr["sub"])
            a.trace_events = _ptr(tr["events"])
            a.trace_k = _ptr(tr["k"])
            a.trace_energy = _ptr(tr["E"])
            if self.tracks_positions:
                tr["pos"] = np.zeros((n_steps, n, 3))
                a.trace_pos = _ptr(tr["pos"])
        self._chk(self.L.emcgpu_step(self.h, C.byref(a)))
        return tr

    # -- positions of bulk particles ---------------------------------------
    def set_box(self, instance, max_pos):
        """emcgpu_set_box: positions tracked in the periodic box [0, max_pos] (m)"""
        self._chk(self.L.emcgpu_set_box(self.h, instance,
                                        (C.c_double * 3)(*[float(x) for x in max_pos])))
        self.tracks_positions = True

    def set_positions(self, instance, pos):
        pos = np.ascontiguousarray(pos, dtype=np.float64).reshape(-1)
        if pos.size != 3 * self.counts[instance]:
            raise ValueError("positions must be [n][3]")
        self._chk(self.L.emcgpu_set_positions(self.h, instance, _ptr(pos)))

    def get_positions(self, instance=0):
        out = np.zeros((self.counts[instance], 3))
        self._chk(self.L.emcgpu_get_positions(self.h, instance, _ptr(out)))
        return out

    def observe(self):
        """emcgpu_observe: the ensemble sums of the current state, (n_instances, n_obs)"""
        n_obs = self.obs_layout()[0]
        out = np.zeros((self.n_instances, n_obs))
        self._chk(self.L.emcgpu_observe(self.h, _ptr(out)))
        return out

    def reserve(self, dt, n_steps, steps_per_launch=0, ppl=0):
        """emcgpu_reserve: size the buffers of a later step() call with these
        arguments (or fewer steps) so that call allocates nothing."""
        a = StepArgs()
        a.dt = dt
        a.n_steps = n_steps
        a.steps_per_launch = steps_per_launch
        a.mode = PRODUCTION
        a.particles_per_lane = ppl
        self._chk(self.L.emcgpu_reserve(self.h, C.byref(a)))

    def step_noalloc(self, dt, n_steps, obs_ptr, obs_len, seed=0, step0=0,
                     steps_per_launch=0, ppl=0, timing=False):
        """emcgpu_step_noalloc: production steps with no device allocation;
        the per-step sums go to the caller's device buffer obs_ptr of
        obs_len doubles, laid out [instance][step][value]."""
        a = StepArgs()
        a.dt = dt
        a.n_steps = n_steps
        a.steps_per_launch = steps_per_launch
        a.mode = PRODUCTION
        a.seed = seed
        a.step0 = step0
        a.timing = 1 if timing else 0
        a.particles_per_lane = ppl
        self._chk(self.L.emcgpu_step_noalloc(self.h, C.byref(a), C.c_void_p(obs_ptr), obs_len))

    def observables(self, instance=0):
        """Raw per-step sums, shape (n_steps, n_obs)."""
        n_obs = C.c_int()
        n_steps = C.c_int()
        self._chk(self.L.emcgpu_get_observables(self.h, instance, None, C.byref(n_obs),
                                                C.byref(n_steps)))
        out = np.zeros(n_obs.value * n_steps.value)
        self._chk(self.L.emcgpu_get_observables(self.h, instance, _ptr(out), C.byref(n_obs),
                                                C.byref(n_steps)))
        return out.reshape(n_steps.value, n_obs.value)

    def obs_layout(self):
        n, mv, md = C.c_int(), C.c_int(), C.c_int()
        self._chk(self.L.emcgpu_obs_layout(self.h, C.byref(n), C.byref(mv), C.byref(md)))
        return n.value, mv.value, md.value

    def averages(self, instance=0):
        """Per-step getAvgEnergy / getAvgDriftVelocity / getValleyOccupationProbability
        (basicBulkParticleHandler.hpp:289-347) plus per-subvalley occupancy."""
        raw = self.observables(instance)
        _, mv, md = self.obs_layout()
        nv = self.n_valleys[instance]
        n = self.counts[instance]
        cnt = raw[:, 2:2 + 3 * mv:3][:, :nv]
        sE = raw[:, 3:3 + 3 * mv:3][:, :nv]
        sV = raw[:, 4:4 + 3 * mv:3][:, :nv]
        with np.errstate(invalid="ignore", divide="ignore"):
            avgE = np.where(cnt > 0, sE / np.maximum(cnt, 1), 0.0)
            avgV = np.where(cnt > 0, sV / np.maximum(cnt, 1), 0.0)
        occ = cnt / max(n, 1)
        sub = raw[:, 2 + 3 * mv:].reshape(raw.shape[0], mv, md)[:, :nv, :] / max(n, 1)
        return dict(avgE=avgE, avgV=avgV, occ=occ, sub_occ=sub, events=raw[:, 0],
                    real_events=raw[:, 1], sumE=sE, sumV=sV, count=cnt)

    def timing(self, reset=False):
        ms = C.c_double()
        n = C.c_longlong()
        self._chk(self.L.emcgpu_get_timing(self.h, C.byref(ms), C.byref(n), 1 if reset else 0))
        return ms.value, n.value

    def synchronize(self):
        self._chk(self.L.emcgpu_synchronize(self.h))

    def autotune(self, dt, seed=0, step0=0, n_steps=0, repeats=0, ppl=(), spl=()):
        """emcgpu_autotune: time every (particles per lane, steps per launch)
        pair on the current ensemble and keep the fastest as the default of
        step(); returns {"best": (ppl, spl, ms_per_step), "trials": [...]}."""
        a = AutotuneArgs()
        a.dt, a.seed, a.step0, a.n_steps, a.repeats = dt, seed, step0, n_steps, repeats
        a.n_ppl = len(ppl)
        for j, v in enumerate(ppl):
            a.ppl[j] = v
        a.n_spl = len(spl)
        for j, v in enumerate(spl):
            a.spl[j] = v
        self._chk(self.L.emcgpu_autotune(self.h, C.byref(a)))
        trials = [(a.res_ppl[j], a.res_spl[j], a.res_ms_per_step[j]) for j in range(a.n_results)]
        return {"best": (a.best_ppl, a.best_spl, a.best_ms_per_step), "trials": trials}

    # -- 2-D device (instance 0) ------------------------------------------
    def device_setup(self, dev, pm=PM_CIC, poisson_tol=1e-10, poisson_max_iter=50,
                     init_from_potential=False, region_tables=False):
        """Geometry from a device dictionary (viennaemc_amd/workloads.py format):
        max_pos, spacing, width, T, eps_r, ni, doping [(min, max, N)],
        contacts [(boundary, voltage, min, max)] for ohmic contacts and
        [(boundary, voltage, min, max, eps_oxide, thickness, barrier_height)]
        for gates."""
        regs = (DopingRegion * max(len(dev["doping"]), 1))()
        for r, (lo, hi, n) in enumerate(dev["doping"]):
            regs[r].min_pos[0], regs[r].min_pos[1] = lo
            regs[r].max_pos[0], regs[r].max_pos[1] = hi
            regs[r].doping = n
        cons = (ContactDesc * max(len(dev["contacts"]), 1))()
        for k, ct in enumerate(dev["contacts"]):
            if len(ct) not in (4, 7):
                raise ValueError("a contact is (boundary, voltage, min, max) or a gate "
                                 "(boundary, voltage, min, max, eps_oxide, thickness, barrier)")
            b, v, lo, hi = ct[:4]
            cons[k].boundary, cons[k].voltage, cons[k].min_pos, cons[k].max_pos = b, v, lo, hi
            if len(ct) == 7:
                cons[k].type = CONTACT_GATE
                cons[k].eps_oxide, cons[k].thickness, cons[k].barrier_height = ct[4:]
        d = DeviceDesc()
        d.max_pos[0], d.max_pos[1] = dev["max_pos"]
        d.spacing[0], d.spacing[1] = dev["spacing"]
        d.width = dev["width"]
        d.temperature = dev["T"]
        d.eps_r = dev["eps_r"]
        d.ni = dev["ni"]
        d.n_doping = len(dev["doping"])
        d.doping = C.cast(regs, C.c_void_p)
        d.n_contacts = len(dev["contacts"])
        d.contacts = C.cast(cons, C.c_void_p)
        d.pm_scheme = pm
        d.poisson_tol = poisson_tol
        d.poisson_max_iter = poisson_max_iter
        d.init_from_potential = 1 if init_from_potential else 0
        d.region_tables = 1 if region_tables else 0
        self._chk(self.L.emcgpu_device_setup(self.h, C.byref(d)))
        self.n_contacts = len(dev["contacts"])
        ext = np.zeros(2, dtype=np.int32)
        sc = np.zeros(4)
        self._chk(self.L.emcgpu_device_info(self.h, _ptr(ext), _ptr(sc)))
        self.ext = (int(ext[0]), int(ext[1]))
        self.nn = self.ext[0] * self.ext[1]
        self.Vt, self.debye, self.cell_volume, self.norm_cell_volume = [float(x) for x in sc]

    def device_grids(self):
        dop = np.zeros(self.nn)
        reg = np.zeros(self.nn, dtype=np.int32)
        con = np.zeros(self.nn, dtype=np.int32)
        self._chk(self.L.emcgpu_device_grids(self.h, _ptr(dop), _ptr(reg), _ptr(con)))
        return dict(doping=dop, region=reg, contact=con)

    def poisson(self, pot, conc=None, equilibrium=False, reset_bc=True):
        pot = np.array(pot, dtype=np.float64).reshape(-1)
        c = None if conc is None else np.ascontiguousarray(conc, dtype=np.float64).reshape(-1)
        _need_size("pot", pot, self.nn)
        if c is not None:
            _need_size("conc", c, self.nn)
        it = C.c_int()
        res = C.c_double()
        self._chk(self.L.emcgpu_poisson(self.h, _ptr(pot), _ptr(c), 1 if equilibrium else 0,
                                        1 if reset_bc else 0, C.byref(it), C.byref(res)))
        return pot, it.value, res.value

    def efield(self, pot):
        pot = np.ascontiguousarray(pot, dtype=np.float64).reshape(-1)
        _need_size("pot", pot, self.nn)
        ex, ey = np.zeros(self.nn), np.zeros(self.nn)
        self._chk(self.L.emcgpu_efield(self.h, _ptr(pot), _ptr(ex), _ptr(ey)))
        return ex, ey

    def deposit(self, pos, grid=None):
        pos = np.ascontiguousarray(pos, dtype=np.float64).reshape(-1, 2)
        g = np.zeros(self.nn) if grid is None else np.array(grid, dtype=np.float64).reshape(-1)
        _need_size("grid", g, self.nn)
        self._chk(self.L.emcgpu_deposit(self.h, _ptr(pos), pos.shape[0], _ptr(g)))
        return g

    def interpolate(self, ex, ey, pos, charge=-Q):
        ex = np.ascontiguousarray(ex, dtype=np.float64).reshape(-1)
        ey = np.ascontiguousarray(ey, dtype=np.float64).reshape(-1)
        _need_size("ex", ex, self.nn)
        _need_size("ey", ey, self.nn)
        pos = np.ascontiguousarray(pos, dtype=np.float64).reshape(-1, 2)
        f = np.zeros((pos.shape[0], 2))
        self._chk(self.L.emcgpu_interpolate(self.h, _ptr(ex), _ptr(ey), _ptr(pos), pos.shape[0],
                                            charge, _ptr(f)))
        return f

    @staticmethod
    def _grid3(ext, spacing, pm):
        g = Grid3()
        for d in range(3):
            g.ext[d] = int(ext[d])
            g.spacing[d] = float(spacing[d])
        g.pm_scheme = int(pm)
        return g

    def deposit3(self, ext, spacing, pos, pm=None, grid=None):
        """3-D assignToMesh (NGP or CIC) of pos[n][3] [m], one carrier each,
        into a grid of ext nodes (x fastest)."""
        g3 = self._grid3(ext, spacing, PM_CIC if pm is None else pm)
        nn = int(np.prod(ext))
        pos = np.ascontiguousarray(pos, dtype=np.float64).reshape(-1, 3)
        g = np.zeros(nn) if grid is None else np.array(grid, dtype=np.float64).reshape(-1)
        _need_size("grid", g, nn)
        self._chk(self.L.emcgpu_deposit3(self.h, C.byref(g3), _ptr(pos), pos.shape[0], _ptr(g)))
        return g

    def interpolate3(self, ext, spacing, ex, ey, ez, pos, charge=-Q, pm=None):
        """3-D interpolateForce: force[n][3] [N] from nodal ex, ey, ez [V/m]."""
        g3 = self._grid3(ext, spacing, PM_CIC if pm is None else pm)
        c = lambda a: np.ascontiguousarray(a, dtype=np.float64).reshape(-1)
        ex, ey, ez = c(ex), c(ey), c(ez)
        for name, a in (("ex", ex), ("ey", ey), ("ez", ez)):
            _need_size(name, a, int(np.prod(ext)))
        pos = np.ascontiguousarray(pos, dtype=np.float64).reshape(-1, 3)
        f = np.zeros((pos.shape[0], 3))
        self._chk(self.L.emcgpu_interpolate3(self.h, C.byref(g3), _ptr(ex), _ptr(ey), _ptr(ez),
                                             _ptr(pos), pos.shape[0], charge, _ptr(f)))
        return f

    def efield3(self, ext, spacing, vt, bc, pot):
        """3-D grid-point E-field (ex, ey, ez) [V/m] from the normalised
        potential; bc[3][nodes] the faces' boundary status (-2 off the
        component's faces, -1 artificial, >= 0 contact)."""
        g3 = self._grid3(ext, spacing, PM_CIC)
        nn = int(np.prod(ext))
        bc = np.ascontiguousarray(bc, dtype=np.int8).reshape(-1)
        pot = np.ascontiguousarray(pot, dtype=np.float64).reshape(-1)
        _need_size("bc", bc, 3 * nn)
        _need_size("pot", pot, nn)
        E = [np.zeros(nn) for _ in range(3)]
        self._chk(self.L.emcgpu_efield3(self.h, C.byref(g3), vt, _ptr(bc), _ptr(pot),
                                        *[_ptr(e) for e in E]))
        return E

    def contacts(self, pos):
        pos = np.ascontiguousarray(pos, dtype=np.float64).reshape(-1, 2)
        keep = np.zeros(pos.shape[0], dtype=np.int32)
        inj = np.zeros(self.nn, dtype=np.int32)
        self._chk(self.L.emcgpu_contacts(self.h, _ptr(pos), pos.shape[0], _ptr(keep), _ptr(inj)))
        return keep, inj

    def device_set_particles(self, pos, k, E, tau, valley, sub):
        pos = np.ascontiguousarray(pos, dtype=np.float64).reshape(-1)
        k = np.ascontiguousarray(k, dtype=np.float64).reshape(-1)
        E = np.ascontiguousarray(E, dtype=np.float64)
        tau = np.ascontiguousarray(tau, dtype=np.float64)
        valley = np.ascontiguousarray(valley, dtype=np.int32)
        sub = np.ascontiguousarray(sub, dtype=np.int32)
        n = E.size
        self._chk(self.L.emcgpu_device_set_particles(self.h, n, _ptr(pos), _ptr(k), _ptr(E),
                                                     _ptr(tau), _ptr(valley), _ptr(sub)))

    def device_get_particles(self):
        n = C.c_size_t(0)
        self._chk(self.L.emcgpu_device_get_particles(self.h, C.byref(n), None, None, None, None,
                                                     None, None))
        m = n.value
        out = dict(pos=np.zeros((m, 2)), k=np.zeros((m, 3)), E=np.zeros(m), tau=np.zeros(m),
                   valley=np.zeros(m, dtype=np.int32), sub=np.zeros(m, dtype=np.int32))
        self._chk(self.L.emcgpu_device_get_particles(
            self.h, C.byref(n), _ptr(out["pos"]), _ptr(out["k"]), _ptr(out["E"]),
            _ptr(out["tau"]), _ptr(out["valley"]), _ptr(out["sub"])))
        return out

    def device_set_efield(self, ex, ey):
        ex = np.ascontiguousarray(ex, dtype=np.float64).reshape(-1)
        ey = np.ascontiguousarray(ey, dtype=np.float64).reshape(-1)
        _need_size("ex", ex, self.nn)
        _need_size("ey", ey, self.nn)
        self._chk(self.L.emcgpu_device_set_efield(self.h, _ptr(ex), _ptr(ey)))

    def device_drift(self, dt, n_steps, n, mode=PRODUCTION, seed=0, step0=0, draws=None,
                     cursor=None, grain_tau=None, trace=False):
        a = DeviceDriftArgs()
        a.dt = dt
        a.n_steps = n_steps
        a.mode = mode
        a.seed = seed
        a.step0 = step0
        keep = []
        if mode == REPLAY:
            draws = np.ascontiguousarray(draws, dtype=np.uint64)
            if cursor.dtype != np.uint64 or not cursor.flags["C_CONTIGUOUS"]:
                raise ValueError("cursor must be a contiguous uint64 array")
            a.draws = _ptr(draws)
            a.draw_stride = draws.shape[1]
            a.cursor = _ptr(cursor)
            if grain_tau is not None:
                a.grain_tau = _ptr(grain_tau)
            keep.append(draws)
        out = dict(exit_step=np.zeros(n, dtype=np.int32), exit_contact=np.zeros(n, dtype=np.int32))
        a.exit_step = _ptr(out["exit_step"])
        a.exit_contact = _ptr(out["exit_contact"])
        if trace:
            out.update(pos=np.zeros((n_steps, n, 2)), k=np.zeros((n_steps, n, 3)),
                       E=np.zeros((n_steps, n)), events=np.zeros((n_steps, n), dtype=np.int32))
            a.tr_pos = _ptr(out["pos"])
            a.tr_k = _ptr(out["k"])
            a.tr_E = _ptr(out["E"])
            a.tr_events = _ptr(out["events"])
        self._chk(self.L.emcgpu_device_drift(self.h, C.byref(a)))
        return out

    def device_init(self, seed):
        n = C.c_size_t(0)
        self._chk(self.L.emcgpu_device_init(self.h, seed, C.byref(n)))
        return n.value

    def device_step(self, dt, n_steps, seed=0, graph=False, poisson_interval=1, channel=None,
                    accumulate=False, poisson_mode=POISSON_AUTO):
        """emcgpu_device_step; channel = (x0, x1, length) [m] records the
        drift-velocity channel current of every step (key "channel_current");
        poisson_mode POISSON_MULTIGRID solves without the Newton corrections from an
        inverse, POISSON_DENSE takes them from the dense inverse even where the
        separable one applies."""
        nc = max(self.n_contacts, 1)
        out = dict(exits=np.zeros((n_steps, nc), dtype=np.int32),
                   deleted=np.zeros((n_steps, nc), dtype=np.int32),
                   injected=np.zeros((n_steps, nc), dtype=np.int32),
                   n_particles=np.zeros(n_steps, dtype=np.int64),
                   poisson_iters=np.zeros(n_steps, dtype=np.int32),
                   poisson_resid=np.zeros(n_steps), checksum=np.zeros(n_steps))
        a = DeviceStepArgs()
        a.dt = dt
        a.n_steps = n_steps
        a.seed = seed
        a.use_graph = 1 if graph else 0
        a.poisson_interval = poisson_interval
        a.accumulate = 1 if accumulate else 0
        a.poisson_mode = poisson_mode
        if channel is not None:
            a.channel_x0, a.channel_x1, a.channel_length = channel
            out["channel_current"] = np.zeros(n_steps)
            a.channel_current = _ptr(out["channel_current"])
        for key in ("exits", "deleted", "injected", "n_particles", "poisson_iters",
                    "poisson_resid", "checksum"):
            setattr(a, key, _ptr(out[key]))
        self._chk(self.L.emcgpu_device_step(self.h, C.byref(a)))
        return out

    def device_averages(self, reset=False):
        """emcgpu_device_averages: (potential, concentration, steps) averaged
        over the steps run with accumulate=True."""
        pot, conc, n = np.zeros(self.nn), np.zeros(self.nn), C.c_longlong(0)
        self._chk(self.L.emcgpu_device_averages(self.h, _ptr(pot), _ptr(conc), C.byref(n),
                                                1 if reset else 0))
        return pot, conc, n.value

    def device_fields(self):
        out = {k: np.zeros(self.nn) for k in ("pot", "ex", "ey", "conc", "nr_part")}
        self._chk(self.L.emcgpu_device_fields(self.h, _ptr(out["pot"]), _ptr(out["ex"]),
                                              _ptr(out["ey"]), _ptr(out["conc"]),
                                              _ptr(out["nr_part"])))
        return out
