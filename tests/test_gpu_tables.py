"""Scatter tables built on the GPU from the mechanism descriptors
(initScatterTables of every instance at once, emcScatterHandler.hpp:220-288)
against the host build of the same C-ABI, whose tables are bit-exact with
the reference's (test_capi_host.py), and against the reference's own rate
tables (golden fixtures).

The rates come from the same FP64 formulas; the device's sqrt is correctly
rounded and its pow / log / exp within an ulp of the host's, and the
compiler may contract a product and a sum into one rounding, so every rate
agrees to a few ulp: rtol 1e-13 on the normalised cumulative tables and on
tau0.  Covered: every kind that the descriptor build takes (acoustic,
zero- and first-order IV in both directions, Froehlich, screened Froehlich),
instances at different temperatures and materials in one launch, and the
kinds it refuses (bath-driven and plasma-screened rates).
"""
import copy

import numpy as np
import pytest

from emc_models import MODELS

pytestmark = pytest.mark.gpu

RTOL = 1e-13
NAMES = ["si_jacoboni", "si_vasileska", "polar_gamma", "para_iso", "two_valley",
         "screened_polar"]


def _emc():
    from viennaemc_amd import emcgpu
    return emcgpu


def _prep(ctx, i, spec):
    ctx.set_model(i, spec)
    scr = spec.get("screening")
    if scr:
        ctx.set_screening(i, scr["qs2"])


@pytest.mark.parametrize("n_levels,max_energy", [(1000, 2.0), (4096, 1.3), (37, 0.9)])
def test_gpu_tables_equal_the_host_build(n_levels, max_energy):
    emcgpu = _emc()
    specs = [MODELS[n] for n in NAMES]
    temps = [300.0 + 37.0 * k for k in range(len(specs))]
    mats = [dict(rho=spec.get("rho", 2329.0), v_sound=spec.get("v_sound", 9040.0))
            for spec in specs]
    mats[0] = dict(rho=2500.0, v_sound=8000.0)
    host = emcgpu.Context(precision=emcgpu.FP64, n_instances=len(specs), device=-1)
    gpu = emcgpu.Context(precision=emcgpu.FP64, n_instances=len(specs), device=0)
    for i, spec in enumerate(specs):
        _prep(host, i, spec)
        _prep(gpu, i, spec)
        host.build_tables(i, n_levels, max_energy, temps[i], rho=mats[i]["rho"],
                          v_sound=mats[i]["v_sound"])
    print(f"build_tables_all n_levels={n_levels} max_energy={max_energy}", flush=True)
    gpu.build_tables_all(n_levels, max_energy, temps, mats)
    for i, spec in enumerate(specs):
        for v in range(len(spec["valleys"])):
            th, tau_h = host.get_tables(i, v, n_levels)
            tg, tau_g = gpu.get_tables(i, v, n_levels)
            assert th.shape == tg.shape
            if th.size == 0:
                continue
            assert np.array_equal(th == 0, tg == 0), (NAMES[i], v)
            nz = th != 0
            assert np.max(np.abs(tg[nz] - th[nz]) / np.abs(th[nz])) <= RTOL, (NAMES[i], v)
            assert abs(tau_g - tau_h) <= RTOL * tau_h, (NAMES[i], v)


@pytest.mark.parametrize("name", NAMES)
def test_gpu_tables_match_the_reference_rates(name, golden):
    """The GPU build against the reference's own rate tables (golden, written
    by the reference's mechanisms at every level, E = (l + 1) dE): each
    valley's cumulative rates over its mechanisms in registration order,
    normalised by the largest total (renormalizeTables,
    emcScatterHandler.hpp:248-272), and tau0 = 1 / that total, within the
    device's few-ulp transcendental differences (rtol 1e-13)."""
    emcgpu = _emc()
    spec = MODELS[name]
    nl, emax, T = spec["n_levels"], spec["max_energy"], spec["T"]
    print("GPU tables vs reference rates:", name, flush=True)
    gpu = emcgpu.Context(precision=emcgpu.FP64, n_instances=1, device=0)
    _prep(gpu, 0, spec)
    gpu.build_tables_all(nl, emax, [T], [dict(rho=spec.get("rho", 2329.0),
                                              v_sound=spec.get("v_sound", 9040.0))])
    rates = golden[name + ".rate_table"].reshape(len(spec["mechs"]), nl)
    mv = golden[name + ".mech_valley"]
    for v in range(len(spec["valleys"])):
        mine = rates[mv == v]
        tg, tau_g = gpu.get_tables(0, v, nl)
        if mine.shape[0] == 0:
            assert tg.size == 0
            continue
        cum = np.cumsum(mine, axis=0)
        top = cum[-1].max()
        ref = cum / top
        nz = ref != 0
        assert np.array_equal(tg == 0, ~nz), (name, v)
        assert np.max(np.abs(tg[nz] - ref[nz]) / ref[nz]) <= RTOL, (name, v)
        assert abs(tau_g - golden[name + ".tau0"][v]) <= RTOL * golden[name + ".tau0"][v]
    gpu.close()


def test_gpu_tables_drive_the_same_step():
    """A short FP64 run from the GPU-built tables equals the run from the
    host-built ones within the table tolerance (same seed, same draws)."""
    emcgpu = _emc()
    spec = MODELS["si_jacoboni"]
    nl, emax, T = spec["n_levels"], spec["max_energy"], spec["T"]
    res = []
    for on_gpu in (False, True):
        ctx = emcgpu.Context(precision=emcgpu.FP64, n_instances=1, device=0)
        ctx.set_model(0, spec)
        if on_gpu:
            ctx.build_tables_all(nl, emax, T, [dict(rho=spec.get("rho", 2329.0),
                                                    v_sound=spec.get("v_sound", 9040.0))])
        else:
            ctx.build_tables(0, nl, emax, T, rho=spec.get("rho", 2329.0),
                             v_sound=spec.get("v_sound", 9040.0))
        ctx.set_field(0, spec["field_dir"], spec["field_kvcm"] * 1e5)
        ctx.resize([4096])
        ctx.init_particles(0, 1234, T)
        ctx.step(dt=1e-14, n_steps=5, seed=99)
        res.append(ctx.get_particles(0))
    k0, E0 = res[0]["k"], res[0]["E"]
    k1, E1 = res[1]["k"], res[1]["E"]
    assert np.array_equal(res[0]["valley"], res[1]["valley"])
    assert np.max(np.abs(E1 - E0) / np.maximum(np.abs(E0), 1e-30)) < 1e-9
    assert np.max(np.abs(k1 - k0)) <= 1e-9 * np.max(np.abs(k0))


@pytest.mark.parametrize("name", ["hot_polar", "screened_hot_polar", "plasma"])
def test_gpu_tables_refuse_dynamic_rates(name):
    emcgpu = _emc()
    ctx = emcgpu.Context(precision=emcgpu.FP64, n_instances=1, device=0)
    if name == "plasma":
        spec = copy.deepcopy(MODELS["screened_polar"])
        ctx.set_model(0, spec)
        ctx.set_screening(0, 0.0, eps_static=12.9, mode=emcgpu.SCREENING_PLASMA, volume=1e-18)
    else:
        spec = MODELS[name]
        ctx.set_model(0, spec)
        ctx.set_baths(0, spec["baths"], spec["T"])
    with pytest.raises(emcgpu.EmcGpuError):
        ctx.build_tables_all(spec["n_levels"], spec["max_energy"], spec["T"])
