"""ctypes binding of the oracle (liboracle.so) -- TEST INFRASTRUCTURE ONLY.

The oracle is the CPU restatement of the reference's bulk EMC step used as
the checker.  Only tests/, __graft_entry__.smoke() and bench.py's
cpu_baseline leg may import this module.
"""
import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None

MAX_DEG = 8  # ORC_MAX_DEG
Q = 1.60219e-19
KB = 1.38066e-23
ME = 9.11e-31

_D = C.c_double
_I = C.c_int
_PD = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")
_PI32 = np.ctypeslib.ndpointer(dtype=np.int32, flags="C_CONTIGUOUS")
_PU64 = np.ctypeslib.ndpointer(dtype=np.uint64, flags="C_CONTIGUOUS")
_PSZ = np.ctypeslib.ndpointer(dtype=np.uintp, flags="C_CONTIGUOUS")


def lib_path():
    return os.path.join(_HERE, "liboracle.so")


def load():
    """Load liboracle.so (built by ``make -C oracle``)."""
    global _LIB
    if _LIB is not None:
        return _LIB
    path = lib_path()
    if not os.path.exists(path):
        raise RuntimeError("oracle not built: run `make -C oracle` (%s missing)" % path)
    L = C.CDLL(path)
    vp = C.c_void_p
    sig = {
        "orc_model_new": (vp, [_I, _D]),
        "orc_model_delete": (None, [vp]),
        "orc_model_add_valley_params": (_I, [vp, _I, _D, _D, _D, _D, _I, _D, _D, C.c_void_p]),
        "orc_model_add_acoustic": (_I, [vp, _I, _D, _D, _D, _D]),
        "orc_model_add_iv": (_I, [vp, _I, _I, _I, _D, _D, _D, _D, _I, _PI32]),
        "orc_model_add_froehlich": (_I, [vp, _I, _I, _D, _D, _D, _D, _D]),
        "orc_model_add_screened_froehlich": (_I, [vp, _I, _I, _D, _D, _D, _D, _D]),
        "orc_model_set_qs2": (None, [vp, _D]),
        "orc_model_add_screened_hot_froehlich": (_I, [vp, _I, _I, _D, _D, _D, _D, _I, _I, _I]),
        "orc_model_set_screening_plasma": (None, [vp, _D, _D]),
        "orc_model_seed_screening": (None, [vp, _I, vp, vp]),
        "orc_model_update_screening": (None, [vp, C.c_longlong, _D, C.c_longlong]),
        "orc_model_get_bath_window": (None, [vp, _I, vp, vp]),
        "orc_model_get_qs2": (_D, [vp]),
        "orc_bath_nq_window_one": (_D, [vp, _I, _D, _D]),
        "orc_bath_sample_q_one": (_D, [vp, _I, _D, _D, _I, _D]),
        "orc_qres_cos_theta_one": (_D, [vp, _I, _D, _D, _I, _D]),
        "orc_scr_log_factor": (_D, [_D, _D, _D]),
        "orc_scr_cos_theta": (_D, [_D, _D, _D, _D]),
        "orc_model_build_tables": (None, [vp]),
        "orc_model_n_tables": (_I, [vp, _I]),
        "orc_model_table_mech": (_I, [vp, _I, _I]),
        "orc_model_tau0": (_D, [vp, _I]),
        "orc_model_dE": (_D, [vp]),
        "orc_model_get_cum": (None, [vp, _I, _PD]),
        "orc_model_rate": (_D, [vp, _I, _D]),
        "orc_energy_level": (_I, [vp, _D]),
        "orc_scatter_one": (_I, [vp, _PD, _PD, _PI32, _PI32, _PU64, C.c_size_t, _PSZ]),
        "orc_mech_scatter_one": (None, [vp, _I, _PD, _PD, _PI32, _PI32, _PU64, C.c_size_t, _PSZ]),
        "orc_drift_one": (None, [vp, _I, _I, _D, _PD, _PD, _PD, _PD]),
        "orc_velocity_one": (None, [vp, _I, _I, _PD, _D, _PD]),
        "orc_normk_one": (_D, [vp, _I, _D]),
        "orc_init_particles_replay": (None, [vp, _I, _D, _PU64, C.c_size_t, _PSZ, _PD, _PD, _PD, _PD, _PI32, _PI32]),
        "orc_run_replay": (_I, [vp, _I, _PD, _PD, _PD, _PD, _PI32, _PI32, _PU64, C.c_size_t, _PSZ, _I, _D, _PD,
                                C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p,
                                C.c_void_p, C.c_void_p]),
        "orc_run_shared": (_I, [vp, C.c_uint64, _PD, _PD, _D, _D, _PD, _D, _D, _I, _I, _PD,
                                C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p,
                                C.c_void_p, C.c_void_p]),
        "orc_run_philox": (None, [vp, _I, _PD, _PD, _PD, _PI32, _PI32, _I, _D, _PD, _PD, C.c_uint64,
                                  C.c_uint32, C.c_uint32, C.c_void_p, C.c_void_p, C.c_void_p]),
        "orc_model_set_box": (None, [vp, _PD]),
        "orc_init_particles_philox": (None, [vp, _I, _D, C.c_uint64, C.c_uint32, _PD, _PD, _PD, _PI32, _PI32,
                                             C.c_void_p]),
        "orc_run_mt": (None, [vp, _I, C.c_uint64, _D, _I, _D, _PD, _PD, C.c_void_p, _I,
                              C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p,
                              C.c_void_p]),
        "orc_init_particles_philox_energy": (None, [vp, _I, _D, _D, C.c_uint64, C.c_uint32, _PD, _PD, _PD,
                                                    _PI32, _PI32]),
        "orc_model_add_bath": (_I, [vp, _I, _D, _D, _D, _D, _D, _I, _D, _D, _D, _D, _D]),
        "orc_model_set_bath_profile": (None, [vp, _I, C.c_void_p]),
        "orc_model_add_hot_froehlich": (_I, [vp, _I, _I, _D, _D, _D, _D, _I]),
        "orc_model_get_bath": (None, [vp, _I, _PD, _PD, _PD, _PD]),
        "orc_model_set_bath": (None, [vp, _I, _PD, _D, _D]),
        "orc_model_bath_record_one": (None, [vp, _I, _D, _I]),
        "orc_model_bath_update_one": (None, [vp, _I, _D]),
        "orc_model_get_bath_energy": (None, [vp, _I, _PD]),
        "orc_model_bath_step": (None, [vp, _D, _I]),
        "orc_model_set_bath_mode": (None, [vp, _I, _I, C.c_longlong]),
        "orc_model_set_init_energy": (None, [vp, _D]),
        "orc_device_new": (vp, [_D, _D, _D, _D, _D, _D, _D]),
        "orc_device_delete": (None, [vp]),
        "orc_device_set_width": (None, [vp, _D]),
        "orc_device_add_doping": (None, [vp, _PD, _PD, _D]),
        "orc_device_add_ohmic": (_I, [vp, _I, _D, _D, _D]),
        "orc_device_add_gate": (_I, [vp, _I, _D, _D, _D, _D, _D, _D]),
        "orc_device_set_region_models": (_I, [vp, C.c_void_p, _I]),
        "orc_device_info": (None, [vp, _PI32, _PD]),
        "orc_device_get_grids": (None, [vp, C.c_void_p, C.c_void_p, C.c_void_p]),
        "orc_sor_solve": (C.c_longlong, [vp, _PD, C.c_void_p, _D, _D, _I]),
        "orc_pm_assign_h": (None, [vp, _I, _PD, _I, _D, _PD]),
        "orc_pm_forces_h": (None, [vp, _I, _PD, _PD, _PD, _I, _D, _PD]),
        "orc_pm_efield": (None, [_I, vp, _PD, _PD, _PD]),
        "orc_pm3_assign": (None, [_I, _PI32, _PD, _PD, _I, _PD]),
        "orc_pm3_force": (None, [_I, _PI32, _PD, _PD, _PD, _PD, _PD, _I, _D, _PD]),
        "orc_pm3_efield": (None, [_PI32, _PD, _D, _PI32, _PD, _PD, _PD, _PD]),
        "orc_device_init_potential": (None, [vp, _PD]),
        "orc_device_concentration": (None, [vp, _PD, _PD]),
        "orc_device_contact_rule": (None, [vp, _PD, _I, _D, _PI32, _PI32, _PD]),
        "orc_device_init_replay": (None, [vp, vp, _I, _PI32, _PU64, C.c_size_t, _PSZ, _PD]),
        "orc_device_run_replay": (None, [vp, vp, _I, _PD, _PD, _I, _PD, _PU64, C.c_size_t, _PSZ, _I, _D,
                                         _PI32, _PI32, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p]),
        "orc_device_run_h": (_I, [vp, vp, _I, _D, _D, _D, _I, C.c_uint64, C.c_void_p, C.c_void_p,
                                  C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p,
                                  C.c_void_p, C.c_void_p, _I, _I, _D, _D, _D, C.c_void_p, _I]),
        "orc_device_run_philox_h": (_I, [vp, vp, _I, _D, _D, _D, _I, C.c_uint64, _D, C.c_void_p,
                                         C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p,
                                         C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, _I,
                                         _I, _D, _D, _D, C.c_void_p, _I]),
        "orc_mt64_fill": (None, [C.c_uint64, _PU64, C.c_size_t]),
        "orc_mt64_fill_particles": (None, [C.c_uint64, C.c_uint64, C.c_size_t, _PU64, C.c_size_t, C.c_size_t]),
        "orc_particle_seed": (C.c_uint64, [C.c_uint64, C.c_uint64]),
        "orc_canonical": (_D, [C.c_uint64]),
        "orc_philox4x32_10": (None, [np.ctypeslib.ndpointer(dtype=np.uint32, flags="C_CONTIGUOUS"),
                                     np.ctypeslib.ndpointer(dtype=np.uint32, flags="C_CONTIGUOUS"),
                                     np.ctypeslib.ndpointer(dtype=np.uint32, flags="C_CONTIGUOUS")]),
        "orc_philox_u01_24": (_D, [C.c_uint32]),
        "orc_prod_uniform": (_D, [C.c_uint64, C.c_uint32, C.c_uint32, C.c_uint32, C.c_uint32]),
    }
    for name, (res, args) in sig.items():
        f = getattr(L, name)
        f.restype = res
        f.argtypes = args
    _LIB = L
    return L


def _ptr(a):
    return None if a is None else a.ctypes.data_as(C.c_void_p)


class OracleModel:
    """A reference particle type (valleys + mechanisms + tables) in the oracle."""

    def __init__(self, spec):
        L = load()
        self.L = L
        self.spec = spec
        self.h = L.orc_model_new(spec["n_levels"], spec["max_energy"])
        T = spec["T"]
        # baths first: the hot-phonon rates read them when the tables are built
        for b in spec.get("baths", []):
            idx = L.orc_model_add_bath(self.h, b["n_bins"], b["dq"], b["tau_lo"], b["hw"], b.get("T", T),
                                       b["volume"], int(b.get("acoustic", False)), b.get("ac_hw", 0.0),
                                       b.get("tau_ac", 0.0), b.get("w_ridley", 0.0), b.get("to_hw", 0.0),
                                       b.get("tau_to", 0.0))
            if b.get("tau_profile") is not None:
                prof = np.ascontiguousarray(b["tau_profile"], dtype=np.float64)
                L.orc_model_set_bath_profile(self.h, idx, prof.ctypes.data_as(C.c_void_p))
        self.n_baths = len(spec.get("baths", []))
        if spec.get("init_energy"):
            L.orc_model_set_init_energy(self.h, spec["init_energy"])
        scr = spec.get("screening")
        if scr:
            L.orc_model_set_qs2(self.h, scr.get("qs2", 0.0))
            if scr.get("plasma"):
                L.orc_model_set_screening_plasma(self.h, scr["eps"], scr["volume"])
        for v in spec["valleys"]:
            axes = None
            if "axes" in v:
                axes = np.ascontiguousarray(np.array(v["axes"], dtype=np.float64).reshape(-1))
            m = v["mass"]
            L.orc_model_add_valley_params(self.h, v["kind"], m[0], m[1], m[2], ME, v["deg"],
                                          v.get("alpha", 0.0), v.get("bottom", 0.0),
                                          None if axes is None else axes.ctypes.data_as(C.c_void_p))
            self._keep = axes
        for mech in spec["mechs"]:
            k = mech["kind"]
            if k == 0:
                L.orc_model_add_acoustic(self.h, mech["valley"], mech["sigma"], spec.get("rho", 2329.0),
                                         spec.get("v_sound", 9040.0), T)
            elif k in (1, 2, 3, 4):
                fmap = np.ascontiguousarray(np.array(mech["fmap"], dtype=np.int32).reshape(-1))
                Z = len(mech["fmap"][0])
                L.orc_model_add_iv(self.h, k, mech["valley"], mech["final"], mech["defpot"], mech["hw"],
                                   spec.get("rho", 2329.0), T, Z, fmap)
            elif k in (7, 8):
                L.orc_model_add_hot_froehlich(self.h, k, mech["valley"], mech["hw"], mech["mrel"],
                                              mech["eps_hi"], mech["eps_lo"], mech["bath"])
            elif k in (11, 12):
                L.orc_model_add_screened_hot_froehlich(
                    self.h, k, mech["valley"], mech["hw"], mech["mrel"], mech["eps_hi"],
                    mech["eps_lo"], mech["bath"], int(mech.get("q_resolved", False)),
                    int(mech.get("q_resolved_angle", True)))
            elif k in (9, 10):
                L.orc_model_add_screened_froehlich(self.h, k, mech["valley"], mech["hw"], mech["mrel"],
                                                   mech["eps_hi"], mech["eps_lo"], T)
            else:
                L.orc_model_add_froehlich(self.h, k, mech["valley"], mech["hw"], mech["mrel"],
                                          mech["eps_hi"], mech["eps_lo"], T)
        L.orc_model_build_tables(self.h)
        self.n_valleys = len(spec["valleys"])

    def __del__(self):
        try:
            self.L.orc_model_delete(self.h)
        except Exception:
            pass

    # -- tables ----------------------------------------------------------
    def n_tables(self, v=0):
        return self.L.orc_model_n_tables(self.h, v)

    def tau0(self, v=0):
        return self.L.orc_model_tau0(self.h, v)

    def cum(self, v=0):
        out = np.zeros(self.n_tables(v) * self.spec["n_levels"], dtype=np.float64)
        self.L.orc_model_get_cum(self.h, v, out)
        return out.reshape(self.n_tables(v), self.spec["n_levels"])

    def rate(self, mech, E):
        return self.L.orc_model_rate(self.h, mech, E)

    def energy_level(self, E):
        return self.L.orc_energy_level(self.h, E)

    # -- single operations ---------------------------------------------
    def drift(self, valley, sub, dt, k, E, pos, force):
        k = np.array(k, dtype=np.float64)
        Ea = np.array([E], dtype=np.float64)
        pos = np.array(pos, dtype=np.float64)
        self.L.orc_drift_one(self.h, valley, sub, dt, k, Ea, pos, np.array(force, dtype=np.float64))
        return k, Ea[0], pos

    def velocity(self, valley, sub, k, E):
        out = np.zeros(3)
        self.L.orc_velocity_one(self.h, valley, sub, np.array(k, dtype=np.float64), E, out)
        return out

    def normk(self, valley, E):
        return self.L.orc_normk_one(self.h, valley, E)

    def mech_scatter(self, mech, k, E, valley, sub, draws):
        k = np.array(k, dtype=np.float64)
        Ea = np.array([E], dtype=np.float64)
        va = np.array([valley], dtype=np.int32)
        sa = np.array([sub], dtype=np.int32)
        used = np.zeros(1, dtype=np.uintp)
        d = np.ascontiguousarray(draws, dtype=np.uint64)
        self.L.orc_mech_scatter_one(self.h, mech, k, Ea, va, sa, d, d.size, used)
        return k, Ea[0], int(va[0]), int(sa[0]), int(used[0])

    # -- ensembles -------------------------------------------------------
    def init_replay(self, draws, cursor, Vt):
        np_ = draws.shape[0]
        k = np.zeros(3 * np_)
        E = np.zeros(np_)
        tau = np.zeros(np_)
        gt = np.zeros(np_)
        v = np.zeros(np_, dtype=np.int32)
        s = np.zeros(np_, dtype=np.int32)
        self.L.orc_init_particles_replay(self.h, np_, Vt, draws, draws.shape[1], cursor, k, E, tau, gt, v, s)
        return dict(k=k, E=E, tau=tau, grain=gt, valley=v, sub=s)

    def run_replay(self, st, draws, cursor, nsteps, dt, force, trace=False):
        """st may hold "pos" [np*3] (m): positions drift and wrap in the box
        (set_box); the trace then holds "pos" per step"""
        np_ = st["E"].size
        outs = {}
        pos = st.get("pos")
        if trace:
            outs = dict(valley=np.zeros(nsteps * np_, dtype=np.int32), sub=np.zeros(nsteps * np_, dtype=np.int32),
                        nev=np.zeros(nsteps * np_, dtype=np.int32), k=np.zeros(nsteps * np_ * 3),
                        E=np.zeros(nsteps * np_))
            if pos is not None:
                outs["pos"] = np.zeros(nsteps * np_ * 3)
        nover = self.L.orc_run_replay(
            self.h, np_, st["k"], st["E"], st["tau"], st["grain"], st["valley"], st["sub"], draws,
            draws.shape[1], cursor, nsteps, dt, np.array(force, dtype=np.float64),
            _ptr(outs.get("valley")), _ptr(outs.get("sub")), _ptr(outs.get("nev")), _ptr(outs.get("k")),
            _ptr(outs.get("E")), None, _ptr(pos), _ptr(outs.get("pos")))
        return nover, outs

    def set_box(self, box):
        self.L.orc_model_set_box(self.h, np.ascontiguousarray(box, dtype=np.float64))

    def run_philox(self, st, nsteps, dt, force, field_dir, seed, instance=0, step0=0, obs=True):
        np_ = st["E"].size
        nv = self.n_valleys
        o = np.zeros(nsteps * 3 * nv) if obs else None
        self.L.orc_run_philox(self.h, np_, st["k"], st["E"], st["tau"], st["valley"], st["sub"], nsteps, dt,
                              np.array(force, dtype=np.float64), np.array(field_dir, dtype=np.float64),
                              seed, instance, step0, _ptr(o), None, _ptr(st.get("pos")))
        return None if o is None else o.reshape(nsteps, 3, nv)

    def init_philox(self, np_, Vt, seed, instance=0, positions=False):
        k = np.zeros(3 * np_)
        E = np.zeros(np_)
        tau = np.zeros(np_)
        v = np.zeros(np_, dtype=np.int32)
        s = np.zeros(np_, dtype=np.int32)
        pos = np.zeros(3 * np_) if positions else None
        self.L.orc_init_particles_philox(self.h, np_, Vt, seed, instance, k, E, tau, v, s, _ptr(pos))
        out = dict(k=k, E=E, tau=tau, valley=v, sub=s)
        if positions:
            out["pos"] = pos
        return out

    def run_mt(self, np_, seed, Vt, nsteps, dt, force, field_dir, obs=True, threads=1, final=False,
               timing=None, sub_occ=False):
        """mt19937_64 ensemble: per-step (avgE, avgV, occ) per valley, shape
        (nsteps, 3, n_valleys); with final the last state; with sub_occ a
        third result, the per-step (valley, subvalley) fractions, shape
        (nsteps, n_valleys, MAX_DEG)."""
        nv = self.n_valleys
        o = np.zeros(nsteps * 3 * nv) if obs else None
        so = np.zeros(nsteps * nv * MAX_DEG) if sub_occ else None
        fin = None
        if final:
            fin = dict(k=np.zeros(3 * np_), E=np.zeros(np_), valley=np.zeros(np_, dtype=np.int32),
                       sub=np.zeros(np_, dtype=np.int32))
        self.L.orc_run_mt(self.h, np_, seed, Vt, nsteps, dt, np.array(force, dtype=np.float64),
                          np.array(field_dir, dtype=np.float64), _ptr(o), threads,
                          _ptr(fin and fin["k"]), _ptr(fin and fin["E"]), _ptr(fin and fin["valley"]),
                          _ptr(fin and fin["sub"]), _ptr(timing), _ptr(so))
        res = ((None if o is None else o.reshape(nsteps, 3, nv)), fin)
        if sub_occ:
            res = res + (so.reshape(nsteps, nv, MAX_DEG),)
        return res

    def run_shared(self, seed, max_pos, spacing, doping, T, field_dir, field_strength, dt, nsteps,
                   max_particles=100000, bath_history=False, state=False):
        """the real handler's loop on one shared stream; with state, a third
        result: the initial positions and the final particles with their
        positions (keys pos0, pos, k, E, valley, sub)"""
        nv = self.n_valleys
        obs = np.zeros((nsteps + 1) * 3 * nv)
        hist = np.zeros(max(nsteps * self.n_baths, 1)) if bath_history else None
        st = None
        if state:
            m = max_particles
            st = dict(pos0=np.zeros(3 * m), pos=np.zeros(3 * m), k=np.zeros(3 * m), E=np.zeros(m),
                      valley=np.zeros(m, dtype=np.int32), sub=np.zeros(m, dtype=np.int32))
        g = (lambda key: _ptr(st[key])) if state else (lambda key: None)
        n = self.L.orc_run_shared(self.h, seed, np.array(max_pos, dtype=np.float64),
                                  np.array(spacing, dtype=np.float64), doping, T,
                                  np.array(field_dir, dtype=np.float64), field_strength, dt, nsteps,
                                  max_particles, obs, g("k"), g("E"), g("valley"), g("sub"),
                                  _ptr(hist), g("pos0"), g("pos"))
        res = (n, obs.reshape(nsteps + 1, 3, nv))
        if bath_history:
            res = res + (hist[:nsteps * self.n_baths].reshape(nsteps, -1),)
        if state:
            for key, per in (("pos0", 3), ("pos", 3), ("k", 3), ("E", 1), ("valley", 1), ("sub", 1)):
                st[key] = st[key][:per * n]
            res = res + (st,)
        return res

    def init_philox_energy(self, np_, E0, sigma, seed, instance=0):
        k = np.zeros(3 * np_)
        E = np.zeros(np_)
        tau = np.zeros(np_)
        v = np.zeros(np_, dtype=np.int32)
        s = np.zeros(np_, dtype=np.int32)
        self.L.orc_init_particles_philox_energy(self.h, np_, E0, sigma, seed, instance, k, E, tau, v, s)
        return dict(k=k, E=E, tau=tau, valley=v, sub=s)

    # -- phonon baths -----------------------------------------------------
    def bath(self, b=0):
        n = self.spec["baths"][b]["n_bins"]
        nq, em, ab, sc = np.zeros(n), np.zeros(n), np.zeros(n), np.zeros(6)
        self.L.orc_model_get_bath(self.h, b, nq, em, ab, sc)
        return dict(nq=nq, n_em=em, n_abs=ab, mean=sc[0], n_ac=sc[1], n_to=sc[2], n0=sc[3],
                    t_ac=sc[4], t_to=sc[5])

    def bath_energy(self, b=0):
        """energy record of the last update [eV]: net emitted phonon energy,
        LO energy gained, LO energy passed to the reservoirs"""
        out = np.zeros(3)
        self.L.orc_model_get_bath_energy(self.h, b, out)
        return out

    def set_bath(self, b, nq, n_ac=0.0, n_to=0.0):
        self.L.orc_model_set_bath(self.h, b, np.ascontiguousarray(nq, dtype=np.float64), n_ac, n_to)

    def bath_record(self, b, q, emission):
        self.L.orc_model_bath_record_one(self.h, b, q, int(emission))

    def bath_update(self, b, dt):
        self.L.orc_model_bath_update_one(self.h, b, dt)

    def qs2(self):
        return self.L.orc_model_get_qs2(self.h)

    def set_qs2(self, qs2):
        """setQs2 on the screening and setScreeningQ2 on every bath."""
        self.L.orc_model_set_qs2(self.h, qs2)

    def seed_screening(self, E, valley):
        """updateScreening + reinitScatterTables before the first step."""
        E = np.ascontiguousarray(E, dtype=np.float64)
        v = np.ascontiguousarray(valley, dtype=np.int32)
        self.L.orc_model_seed_screening(self.h, E.size, _ptr(E), _ptr(v))

    def bath_window(self, b=0):
        nb = self.spec["baths"][b]["n_bins"]
        w, wn = np.zeros(nb + 1), np.zeros(nb + 1)
        self.L.orc_model_get_bath_window(self.h, b, _ptr(w), _ptr(wn))
        return w, wn

    def nq_in_window(self, b, qmin, qmax):
        return self.L.orc_bath_nq_window_one(self.h, b, qmin, qmax)

    def sample_q(self, b, qmin, qmax, emission, r):
        return self.L.orc_bath_sample_q_one(self.h, b, qmin, qmax, int(emission), r)

    def qres_cos_theta(self, b, kI, kF, emission, r):
        return self.L.orc_qres_cos_theta_one(self.h, b, kI, kF, int(emission), r)

    def bath_step(self, dt, rebuild=True):
        self.L.orc_model_bath_step(self.h, dt, int(rebuild))

    def set_bath_mode(self, update=True, rebuild_every=1, step_count=0):
        self.L.orc_model_set_bath_mode(self.h, int(update), rebuild_every, step_count)


def mt64_fill(seed, n):
    out = np.zeros(n, dtype=np.uint64)
    load().orc_mt64_fill(seed, out, n)
    return out


def mt64_particle_draws(seed, np_, n, p0=0):
    out = np.zeros((np_, n), dtype=np.uint64)
    load().orc_mt64_fill_particles(seed, p0, np_, out, n, n)
    return out


def particle_seed(seed, pid):
    return load().orc_particle_seed(seed, pid)


def canonical(x):
    return load().orc_canonical(x)


def philox(ctr, key):
    out = np.zeros(4, dtype=np.uint32)
    load().orc_philox4x32_10(np.array(ctr, dtype=np.uint32), np.array(key, dtype=np.uint32), out)
    return out


def prod_uniform(seed, inst, pid, step, j):
    """uniform j of the production (step, particle, instance) stream"""
    return load().orc_prod_uniform(seed, inst, pid, step, j)


def pm3_assign(scheme, ext, h, pos):
    """3-D assignToMesh (NGP or CIC) of pos[n][3], one carrier each:
    the grid [ext[2]][ext[1]][ext[0]] flattened x fastest."""
    L = load()
    ext = np.ascontiguousarray(ext, dtype=np.int32)
    pos = np.ascontiguousarray(pos, dtype=np.float64).reshape(-1)
    g = np.zeros(int(np.prod(ext)))
    L.orc_pm3_assign(scheme, ext, np.ascontiguousarray(h, dtype=np.float64), pos, pos.size // 3, g)
    return g


def pm3_force(scheme, ext, h, ex, ey, ez, pos, charge):
    """3-D interpolateForce of pos[n][3]: force[n][3]."""
    L = load()
    ext = np.ascontiguousarray(ext, dtype=np.int32)
    pos = np.ascontiguousarray(pos, dtype=np.float64).reshape(-1)
    f = np.zeros(pos.size)
    c = lambda a: np.ascontiguousarray(a, dtype=np.float64)
    L.orc_pm3_force(scheme, ext, c(h), c(ex), c(ey), c(ez), pos, pos.size // 3, charge, f)
    return f.reshape(-1, 3)


def pm3_efield(ext, h, vt, bc, pot):
    """3-D calcEFieldAtGridPts: (ex, ey, ez) from the normalised potential,
    bc[3][nodes] as in orc_pm3_efield."""
    L = load()
    ext = np.ascontiguousarray(ext, dtype=np.int32)
    nn = int(np.prod(ext))
    E = [np.zeros(nn) for _ in range(3)]
    L.orc_pm3_efield(ext, np.ascontiguousarray(h, dtype=np.float64), vt,
                     np.ascontiguousarray(bc, dtype=np.int32).reshape(-1),
                     np.ascontiguousarray(pot, dtype=np.float64), *E)
    return E


def thermal_voltage(T):
    """kB / q * T (emcDevice.hpp:88)."""
    return KB / Q * T


PM_NGP, PM_CIC, PM_NEC = 0, 1, 2


class OracleDevice:
    """A reference 2-D device (emcDevice + doping + contacts) in the oracle.
    A contact is (boundary, voltage, min, max) for an ohmic contact or
    (boundary, voltage, min, max, eps_oxide, thickness, barrier_height) for a
    gate (addGateContact, emcDevice.hpp:195-202)."""

    def __init__(self, spec):
        L = load()
        self.L = L
        self.spec = spec
        mx, my = spec["max_pos"]
        hx, hy = spec["spacing"]
        self.h = L.orc_device_new(mx, my, hx, hy, spec["T"], spec["eps_r"], spec["ni"])
        L.orc_device_set_width(self.h, spec["width"])
        for lo, hi, dop in spec["doping"]:
            L.orc_device_add_doping(self.h, np.array(lo, dtype=np.float64), np.array(hi, dtype=np.float64),
                                    dop)
        for ct in spec["contacts"]:
            if len(ct) == 4:
                ok = L.orc_device_add_ohmic(self.h, *ct) >= 0
            else:
                ok = L.orc_device_add_gate(self.h, *ct) >= 0
            if not ok:
                raise ValueError("invalid contact")
        ext = np.zeros(2, dtype=np.int32)
        sc = np.zeros(4)
        L.orc_device_info(self.h, ext, sc)
        self.ext = (int(ext[0]), int(ext[1]))
        self.n_nodes = self.ext[0] * self.ext[1]
        self.Vt, self.debye, self.cell_volume, self.norm_cell_volume = sc

    def __del__(self):
        try:
            self.L.orc_device_delete(self.h)
        except Exception:
            pass

    def set_region_models(self, models):
        """per doping region the particle type's tables of that region
        (OracleModel list; [] clears): scattering and new free-flight times
        of a particle then use the model of its region"""
        self._region_models = list(models)  # keep the handles alive
        arr = (C.c_void_p * max(len(models), 1))(*[m.h for m in models])
        if self.L.orc_device_set_region_models(self.h, C.cast(arr, C.c_void_p), len(models)) != 0:
            raise ValueError("too many region models")

    def grids(self):
        d = np.zeros(self.n_nodes)
        r = np.zeros(self.n_nodes, dtype=np.int32)
        o = np.zeros(self.n_nodes, dtype=np.int32)
        self.L.orc_device_get_grids(self.h, _ptr(d), _ptr(r), _ptr(o))
        return dict(doping=d, region=r, contact=o)

    def init_potential(self):
        pot = np.zeros(self.n_nodes)
        self.L.orc_device_init_potential(self.h, pot)
        return pot

    def sor(self, pot, conc=None, accuracy=1e-12, omega=1.8, reset_bc=True):
        pot = np.array(pot, dtype=np.float64)
        c = None if conc is None else np.ascontiguousarray(conc, dtype=np.float64)
        sweeps = self.L.orc_sor_solve(self.h, pot, _ptr(c), accuracy, omega, int(reset_bc))
        return pot, sweeps

    def assign(self, scheme, pos, carriers=1.0):
        pos = np.ascontiguousarray(pos, dtype=np.float64).reshape(-1)
        g = np.zeros(self.n_nodes)
        self.L.orc_pm_assign_h(self.h, scheme, pos, pos.size // 2, carriers, g)
        return g

    def forces(self, scheme, ex, ey, pos, charge=-Q):
        pos = np.ascontiguousarray(pos, dtype=np.float64).reshape(-1)
        f = np.zeros(3 * (pos.size // 2))
        self.L.orc_pm_forces_h(self.h, scheme, np.ascontiguousarray(ex, dtype=np.float64),
                               np.ascontiguousarray(ey, dtype=np.float64), pos, pos.size // 2, charge, f)
        return f.reshape(-1, 3)

    def efield(self, scheme, pot):
        ex, ey = np.zeros(self.n_nodes), np.zeros(self.n_nodes)
        self.L.orc_pm_efield(scheme, self.h, np.ascontiguousarray(pot, dtype=np.float64), ex, ey)
        return ex, ey

    def concentration(self, nr_part):
        c = np.zeros(self.n_nodes)
        self.L.orc_device_concentration(self.h, np.ascontiguousarray(nr_part, dtype=np.float64), c)
        return c

    def contact_rule(self, pos, carriers=1.0):
        pos = np.ascontiguousarray(pos, dtype=np.float64).reshape(-1)
        n = pos.size // 2
        keep = np.zeros(max(n, 1), dtype=np.int32)
        inj = np.zeros(self.n_nodes, dtype=np.int32)
        exp = np.zeros(self.n_nodes)
        self.L.orc_device_contact_rule(self.h, pos, n, carriers, keep, inj, exp)
        return keep[:n], inj, exp

    def init_replay(self, model, coords, draws, cursor):
        st = np.zeros((len(coords), 10))
        self.L.orc_device_init_replay(model.h, self.h, len(coords),
                                      np.ascontiguousarray(coords, dtype=np.int32).reshape(-1),
                                      draws, draws.shape[1], cursor, st)
        return st

    def run_replay(self, model, scheme, ex, ey, state, draws, cursor, nsteps, dt, trace=True):
        np_ = state.shape[0]
        rem = np.full(np_, -1, dtype=np.int32)
        con = np.full(np_, -1, dtype=np.int32)
        tr = {}
        if trace:
            tr = dict(pos=np.zeros(nsteps * np_ * 2), k=np.zeros(nsteps * np_ * 3), E=np.zeros(nsteps * np_),
                      nev=np.zeros(nsteps * np_, dtype=np.int32))
        self.L.orc_device_run_replay(model.h, self.h, scheme, np.ascontiguousarray(ex, dtype=np.float64),
                                     np.ascontiguousarray(ey, dtype=np.float64), np_, state, draws,
                                     draws.shape[1], cursor, nsteps, dt, rem, con, _ptr(tr.get("pos")),
                                     _ptr(tr.get("k")), _ptr(tr.get("E")), _ptr(tr.get("nev")))
        return rem, con, tr

    def run(self, model, scheme, accuracy, omega, dt, nsteps, seed, max_state=0, interval=1,
            channel=None, init_from_potential=False):
        """the emcSimulation step order with the reference engine's streams;
        interval: setPoissonInterval, channel = (x0, x1, length) [m]: the
        drift-velocity channel current of every step (key "chan")"""
        nn, nc = self.n_nodes, len(self.spec["contacts"])
        out = dict(eq_pot=np.zeros(nn), n_init=np.zeros(1, dtype=np.int32),
                   n_part=np.zeros(nsteps, dtype=np.int32), pot=np.zeros(nsteps * nn),
                   exits=np.zeros(nsteps * nc, dtype=np.int32), dels=np.zeros(nsteps * nc, dtype=np.int32),
                   injs=np.zeros(nsteps * nc, dtype=np.int32), nr_part=np.zeros(nn),
                   state=np.zeros(max(max_state, 1) * 9))
        ch = channel or (0.0, 0.0, 0.0)
        out["chan"] = np.zeros(nsteps) if channel else None
        n = self.L.orc_device_run_h(self.h, model.h, scheme, accuracy, omega, dt, nsteps, seed,
                                    _ptr(out["eq_pot"]), _ptr(out["n_init"]), _ptr(out["n_part"]),
                                    _ptr(out["pot"]), _ptr(out["exits"]), _ptr(out["dels"]),
                                    _ptr(out["injs"]), _ptr(out["nr_part"]), _ptr(out["state"]), max_state,
                                    interval, ch[0], ch[1], ch[2], _ptr(out["chan"]),
                                    1 if init_from_potential else 0)
        out["n_final"] = n
        out["pot"] = out["pot"].reshape(nsteps, nn)
        for k in ("exits", "dels", "injs"):
            out[k] = out[k].reshape(nsteps, nc)
        out["state"] = out["state"][:min(n, max_state) * 9].reshape(-1, 9)
        return out

    def run_philox(self, model, scheme, accuracy, omega, dt, nsteps, seed, Vt, max_state=0,
                   interval=1, channel=None, init_from_potential=False):
        """the production twin of the GPU device run (Philox streams); state
        rows: x, y, kx, ky, kz, E, tau, valley, sub, pid; interval and
        channel as in run"""
        nn, nc = self.n_nodes, len(self.spec["contacts"])
        out = dict(eq_pot=np.zeros(nn), n_init=np.zeros(1, dtype=np.int32),
                   n_part=np.zeros(nsteps, dtype=np.int32), pot=np.zeros(nsteps * nn),
                   exits=np.zeros(nsteps * nc, dtype=np.int32), dels=np.zeros(nsteps * nc, dtype=np.int32),
                   injs=np.zeros(nsteps * nc, dtype=np.int32), nr_part=np.zeros(nn),
                   state=np.zeros(max(max_state, 1) * 10))
        ch = channel or (0.0, 0.0, 0.0)
        out["chan"] = np.zeros(nsteps) if channel else None
        n = self.L.orc_device_run_philox_h(
            self.h, model.h, scheme, accuracy, omega, dt, nsteps, seed, Vt, _ptr(out["eq_pot"]),
            _ptr(out["n_init"]), _ptr(out["n_part"]), _ptr(out["pot"]), _ptr(out["exits"]),
            _ptr(out["dels"]), _ptr(out["injs"]), _ptr(out["nr_part"]), _ptr(out["state"]),
            max_state, interval, ch[0], ch[1], ch[2], _ptr(out["chan"]),
            1 if init_from_potential else 0)
        out["n_final"] = n
        out["pot"] = out["pot"].reshape(nsteps, nn)
        for k in ("exits", "dels", "injs"):
            out[k] = out[k].reshape(nsteps, nc)
        out["state"] = out["state"][:min(n, max_state) * 10].reshape(-1, 10)
        return out
