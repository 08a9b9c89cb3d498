"""C-ABI library checks that need no GPU.

libemcgpu.so loads, exports every entry point declared in include/emcgpu.h,
and its host-side table construction (host-only context, device = -1)
reproduces the reference's FP64 tables; invalid models are rejected with
the reference's error conditions.
"""
import ctypes as C
import os
import re
import subprocess

import numpy as np
import pytest

from emc_models import MODELS
from oracle import oracle_py as O
from viennaemc_amd import emcgpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    names = set()
    for hdr in os.listdir(os.path.join(ROOT, "include")):
        if hdr.endswith(".h"):
            src = open(os.path.join(ROOT, "include", hdr)).read()
            names |= set(re.findall(r"^\s*(?:const\s+)?\w+\s*\*?\s*(emcgpu_\w+)\s*\(", src, re.M))
    return names


def test_library_exports_every_declared_symbol():
    lib = emcgpu.LIB_PATH
    assert os.path.exists(lib), "build libemcgpu.so first"
    out = subprocess.check_output(["nm", "-D", "--defined-only", lib]).decode()
    exported = set(re.findall(r" T (emcgpu_\w+)", out))
    decl = declared_symbols()
    assert len(decl) >= 18
    assert decl <= exported, decl - exported
    L = emcgpu.load()
    for name in decl:
        assert hasattr(L, name)
    assert set(emcgpu.EXPORTS) <= decl


def test_struct_layouts_match_header():
    # sizes computed by the C compiler for the descriptors of include/emcgpu.h
    assert C.sizeof(emcgpu.ValleyDesc) == 4 + 4 + 24 + 8 + 4 + 4 + 8 + 8 + 8 * 9 * 8
    assert C.sizeof(emcgpu.MechDesc) == 3 * 4 + 4 + 5 * 8 + 4 + 8 * 8 * 4 + 4 + 2 * 4
    assert emcgpu.MechDesc.bath.offset == 316
    assert emcgpu.MechDesc.q_resolved_angle.offset == 324
    assert C.sizeof(emcgpu.ScreeningDesc) == 4 + 4 + 3 * 8
    assert C.sizeof(emcgpu.BathDesc) == 4 + 4 + 5 * 8 + 4 + 4 + 5 * 8 + 8
    assert C.sizeof(emcgpu.StepArgs) == 8 + 4 * 3 + 4 + 8 + 4 + 4 + 4 + 4 + 4 + 4 + 8 * 10
    assert emcgpu.StepArgs.draws.offset == 56


def test_step_arg_layouts_match_the_c_compiler(tmp_path):
    """sizeof / offsetof of the step-argument, contact and device structs
    from gcc on the header itself, against the ctypes mirrors of the
    binding"""
    src = tmp_path / "layout.c"
    src.write_text(
        '#include <stdio.h>\n#include <stddef.h>\n#include "emcgpu.h"\n'
        'int main(void) {\n'
        '  printf("%zu %zu %zu\\n", sizeof(emcgpu_step_args), sizeof(emcgpu_device_step_args),\n'
        '         offsetof(emcgpu_device_step_args, poisson_mode));\n'
        '  printf("%zu %zu\\n", offsetof(emcgpu_step_args, trace_pos),\n'
        '         offsetof(emcgpu_device_step_args, accumulate));\n'
        '  printf("%zu %zu %zu\\n", sizeof(emcgpu_contact), offsetof(emcgpu_contact, type),\n'
        '         offsetof(emcgpu_contact, barrier_height));\n'
        '  printf("%zu %zu\\n", sizeof(emcgpu_device_desc),\n'
        '         offsetof(emcgpu_device_desc, init_from_potential));\n'
        '  return 0;\n}\n')
    exe = tmp_path / "layout"
    inc = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "include")
    subprocess.check_call(["gcc", "-std=c11", "-I", inc, str(src), "-o", str(exe)])
    a, b, c, d = subprocess.check_output([str(exe)]).decode().split("\n")[:4]
    s_step, s_dev, o_mode = (int(x) for x in a.split())
    o_trace, o_acc = (int(x) for x in b.split())
    assert C.sizeof(emcgpu.StepArgs) == s_step
    assert emcgpu.StepArgs.trace_pos.offset == o_trace
    assert C.sizeof(emcgpu.DeviceStepArgs) == s_dev
    assert emcgpu.DeviceStepArgs.poisson_mode.offset == o_mode
    assert emcgpu.DeviceStepArgs.accumulate.offset == o_acc
    s_con, o_type, o_bar = (int(x) for x in c.split())
    assert C.sizeof(emcgpu.ContactDesc) == s_con
    assert emcgpu.ContactDesc.type.offset == o_type
    assert emcgpu.ContactDesc.barrier_height.offset == o_bar
    s_desc, o_init = (int(x) for x in d.split())
    assert C.sizeof(emcgpu.DeviceDesc) == s_desc
    assert emcgpu.DeviceDesc.init_from_potential.offset == o_init


def host_ctx(n_instances=1):
    return emcgpu.Context(precision=emcgpu.FP64, n_instances=n_instances, device=-1)


@pytest.mark.parametrize("name", list(MODELS))
def test_host_tables_match_reference(name, golden, oracle_models):
    spec = MODELS[name]
    ctx = host_ctx()
    ctx.set_model(0, spec)
    if spec.get("baths"):
        ctx.set_baths(0, spec["baths"], spec["T"])
    if spec.get("screening"):
        ctx.set_screening(0, spec["screening"]["qs2"])
    ctx.build_tables(0, spec["n_levels"], spec["max_energy"], spec["T"])
    M = O.OracleModel(spec) if spec.get("baths") else oracle_models[name]
    for v in range(len(spec["valleys"])):
        cum, tau0 = ctx.get_tables(0, v, spec["n_levels"])
        # tau0 is the reference's getTau; the tables are the reference's
        # normalised cumulative tables (bit-exact FP64)
        assert tau0 == golden[name + ".tau0"][v]
        assert (cum == M.cum(v)).all()


def test_user_rate_tables_match_builtin(golden):
    """emcgpu_set_rate_tables with the reference's raw rates == built-in tables."""
    name = "si_jacoboni"
    spec = MODELS[name]
    nm = len(spec["mechs"])
    raw = golden[name + ".rate_table"].reshape(nm, -1)
    a = host_ctx()
    a.set_model(0, spec)
    a.set_rate_tables(0, spec["max_energy"], raw)
    b = host_ctx()
    b.set_model(0, spec)
    b.build_tables(0, spec["n_levels"], spec["max_energy"], spec["T"])
    ca, ta = a.get_tables(0, 0, spec["n_levels"])
    cb, tb = b.get_tables(0, 0, spec["n_levels"])
    assert ta == tb and (ca == cb).all()


def _bad(spec_mut):
    import copy
    spec = copy.deepcopy(MODELS["si_jacoboni"])
    spec_mut(spec)
    ctx = host_ctx()
    with pytest.raises(emcgpu.EmcGpuError) as e:
        ctx.set_model(0, spec)
    return str(e.value)


def test_rejects_non_orthogonal_axes():
    def m(s):
        s["valleys"][0]["axes"][1] = [(1, 1, 0), (1, 0, 0), (0, 0, 1)]
    assert "not orthogonal" in _bad(m)


def test_rejects_bad_final_subvalley():
    def m(s):
        s["mechs"][1]["fmap"] = [[3], [1], [2]]
    assert "not valid" in _bad(m)


def test_rejects_negative_nonparabolicity():
    def m(s):
        s["valleys"][0]["alpha"] = -0.1
    assert "NonParabolicity" in _bad(m)


def test_rejects_bad_valley_index():
    def m(s):
        s["mechs"][0]["valley"] = 3
    assert "invalid" in _bad(m)


def test_host_context_refuses_device_work():
    ctx = host_ctx()
    with pytest.raises(emcgpu.EmcGpuError) as e:
        ctx.resize([10])
    assert e.value.status == 3


def test_tables_before_model_is_a_state_error():
    ctx = host_ctx()
    with pytest.raises(emcgpu.EmcGpuError) as e:
        ctx.build_tables(0, 100, 1.0, 300.0)
    assert e.value.status == 3


def test_rejects_negative_user_rates(golden):
    spec = MODELS["si_jacoboni"]
    raw = golden["si_jacoboni.rate_table"].reshape(len(spec["mechs"]), -1).copy()
    raw[3, 10] = -1.0
    ctx = host_ctx()
    ctx.set_model(0, spec)
    with pytest.raises(emcgpu.EmcGpuError) as e:
        ctx.set_rate_tables(0, spec["max_energy"], raw)
    assert "non-negative" in str(e.value)


# ---- phonon baths (host-only context) ----------------------------------------

def _hot_ctx(spec=None):
    spec = spec or MODELS["hot_polar"]
    ctx = host_ctx()
    ctx.set_model(0, spec)
    ctx.set_baths(0, spec["baths"], spec["T"])
    return ctx, spec


def test_bath_initial_state_matches_oracle():
    ctx, spec = _hot_ctx()
    M = O.OracleModel(spec)
    got, ref = ctx.get_bath(0, 0), M.bath(0)
    assert (got["nq"] == ref["nq"]).all()
    for key in ("mean", "n_ac", "n_to", "n0", "t_ac", "t_to"):
        assert got[key] == ref[key], key


def test_bath_state_drives_host_tables():
    """A hot bath state set before the table build gives the oracle's tables
    for the same state (rates read getMeanNq)."""
    ctx, spec = _hot_ctx()
    nb = spec["baths"][0]["n_bins"]
    rng = np.random.default_rng(5)
    nq = rng.uniform(0.0, 3.0, nb)
    ctx.set_bath_state(0, 0, nq, 0.0, 0.0)
    M2 = O.OracleModel(spec)
    M2.set_bath(0, nq)
    M2.L.orc_model_build_tables(M2.h)
    ctx.build_tables(0, spec["n_levels"], spec["max_energy"], spec["T"])
    cum, tau0 = ctx.get_tables(0, 0, spec["n_levels"])
    assert tau0 == M2.tau0(0)
    assert (cum == M2.cum(0)).all()
    assert ctx.get_bath(0, 0)["mean"] == M2.bath(0)["mean"]


def test_bath_validation():
    spec = MODELS["hot_polar"]
    ctx = host_ctx()
    ctx.set_model(0, spec)
    with pytest.raises(emcgpu.EmcGpuError, match="set the baths first"):
        ctx.build_tables(0, spec["n_levels"], spec["max_energy"], spec["T"])
    bad = dict(spec["baths"][0], n_bins=0)
    with pytest.raises(emcgpu.EmcGpuError, match="n_bins"):
        ctx.set_baths(0, [bad], 300.0)
    bad = dict(spec["baths"][0], volume=0.0)
    with pytest.raises(emcgpu.EmcGpuError, match="must be > 0"):
        ctx.set_baths(0, [bad], 300.0)
    bad = dict(spec["baths"][0], acoustic=True, ac_hw=0.03, tau_ac=0.0)
    with pytest.raises(emcgpu.EmcGpuError, match="acoustic"):
        ctx.set_baths(0, [bad], 300.0)
    ctx.set_baths(0, spec["baths"], 300.0)
    with pytest.raises(emcgpu.EmcGpuError, match="follow their bath"):
        ctx.set_rate_tables(0, spec["max_energy"], np.ones((2, spec["n_levels"])))
    with pytest.raises(emcgpu.EmcGpuError, match="occupations"):
        ctx.set_bath_state(0, 0, -np.ones(spec["baths"][0]["n_bins"]))
    # a mechanism pointing past the supported baths
    m2 = dict(spec, mechs=[dict(spec["mechs"][0], bath=7)])
    with pytest.raises(emcgpu.EmcGpuError, match="bath index"):
        ctx.set_model(0, m2)


def test_build_tables_all_needs_a_device_and_valid_arguments():
    """emcgpu_build_tables_all evaluates the rates on the GPU: a host-only
    context, a bad level count or energy, and missing arrays are refused
    before any work (no compute call without a GPU)."""
    from viennaemc_amd import emcgpu
    ctx = emcgpu.Context(precision=emcgpu.FP64, n_instances=2, device=-1)
    ctx.set_model(0, MODELS["si_jacoboni"])
    ctx.set_model(1, MODELS["polar_gamma"])
    with pytest.raises(emcgpu.EmcGpuError, match="needs a device"):
        ctx.build_tables_all(1000, 2.0, 300.0)
    L = ctx.L
    assert L.emcgpu_build_tables_all(None, 10, 1.0, None, None) != 0
