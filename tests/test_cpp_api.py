"""The C++17 header-only host API (viennaemc_amd/include/emcgpu/).

CPU: the reference's bulk-silicon particle type written against the host API
lowers to the same scatter rates and tau0 as the reference (golden values,
bit-exact) and keeps the reference's subvalley-frame validation.  The
result-file writers of device simulations reproduce, byte for byte, the
files the reference's emcSimulation wrote (grids, contact currents,
particle dumps) from the numbers in them, and the particle type's
scatter-table dumps equal the reference's *ScatterMechanism.txt files (bulk
silicon and the device type in three doping regions).
GPU: the C++ example (bulkSimulationGpu, the reference's bulkSimulation
written against the host API) produces exactly the observables of the
Python binding on the same seed.
"""
import os
import subprocess

import numpy as np
import pytest

from emc_models import MODELS

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG = os.path.join(ROOT, "viennaemc_amd")


@pytest.fixture(scope="module")
def host_api_binary(tmp_path_factory):
    out = str(tmp_path_factory.mktemp("cpp") / "host_api_tables")
    subprocess.check_call(["g++", "-std=c++17", "-O2", "-I", os.path.join(PKG, "include"),
                           os.path.join(ROOT, "tests", "cpp", "host_api_tables.cpp"), "-o", out,
                           "-L", PKG, "-lemcgpu", "-Wl,-rpath," + PKG])
    return out


def test_scatter_table_dumps_match_the_reference(host_api_binary, golden, tmp_path):
    subprocess.check_call([host_api_binary, str(tmp_path)], stdout=subprocess.DEVNULL)
    keys = [k for k in golden.files if k.startswith("tabfile.")]
    assert len(keys) == 18
    for k in keys:
        _, kind, name = k.split(".")
        ours = (tmp_path / ("%s_%s.txt" % ("si" if kind == "si_vasileska" else kind, name))).read_bytes()
        assert ours == bytes(golden[k]), k


def test_host_api_tables_match_reference(host_api_binary, golden):
    txt = subprocess.check_output([host_api_binary]).decode().splitlines()
    rates = np.array([float(l.split()[1]) for l in txt if l.startswith("rate ")])
    ref = golden["si_vasileska.rate_table"]
    assert rates.shape == ref.shape and (rates == ref).all()
    tau0 = float([l for l in txt if l.startswith("tau0")][0].split()[1])
    assert tau0 == golden["si_vasileska.tau0"][0]
    assert any(l.startswith("rejected") and "not orthogonal" in l for l in txt)


@pytest.mark.gpu
def test_cpp_example_matches_python_binding():
    from viennaemc_amd import emcgpu
    exe = os.path.join(PKG, "examples", "bulkSimulationGpu")
    n, steps, seed = 20000, 100, 4321
    print("cpp example", n, steps, flush=True)
    out = subprocess.check_output([exe, str(n), str(steps), "fp64", str(seed), "50"],
                                  env=dict(os.environ, EMCGPU_THROW="1"), timeout=120).decode()
    rows = [l.split() for l in out.splitlines() if l.startswith("step ")]
    spec = MODELS["si_vasileska"]
    ctx = emcgpu.Context(precision=emcgpu.FP64)
    ctx.setup(0, spec)
    ctx.resize([n])
    ctx.init_particles(0, seed, spec["T"])
    ctx.step(1e-15, steps, seed=seed)
    av = ctx.averages(0)
    for r in rows:
        s = int(r[1]) - 1
        assert float(r[2]) == av["avgE"][s, 0]
        assert float(r[3]) == av["avgV"][s, 0]
        assert float(r[4]) == av["occ"][s, 0]


@pytest.mark.gpu
def test_cpp_example_files_match_python_binding(tmp_path):
    """bulkSimulationGpu with a file prefix: the bulk example's five files
    (particle dumps with positions after initialisation and at the end, and
    the per-step time series from step 0) hold exactly the state and the
    averages of the Python binding on the same seed, in the reference's
    text format."""
    from viennaemc_amd import emcgpu
    exe = os.path.join(PKG, "examples", "bulkSimulationGpu")
    n, steps, seed, every, dt = 5000, 60, 777, 25, 1e-15
    box = [5e-7] * 3
    print("cpp example files", n, steps, flush=True)
    prefix = str(tmp_path / "bulkSimulation")
    subprocess.check_call([exe, str(n), str(steps), "fp64", str(seed), str(every), prefix],
                          env=dict(os.environ, EMCGPU_THROW="1"), timeout=120,
                          stdout=subprocess.DEVNULL)
    spec = MODELS["si_vasileska"]
    ctx = emcgpu.Context(precision=emcgpu.FP64)
    ctx.setup(0, spec)
    ctx.resize([n])
    ctx.set_box(0, box)
    ctx.init_particles(0, seed, spec["T"])

    def dump():
        g, pos = ctx.get_particles(0), ctx.get_positions(0)
        k = g["k"].reshape(-1, 3)
        lines = [" ".join("%g" % x for x in box)]
        lines += ["%d %s %s %g %d %d" % (i, " ".join("%g" % x for x in pos[i]),
                                         " ".join("%g" % x for x in k[i]), g["E"][i],
                                         g["sub"][i], g["valley"][i]) for i in range(n)]
        return "\n".join(lines).encode()

    def read(part):
        return (tmp_path / ("bulkSimulation%s.txt" % part)).read_bytes()

    assert read("ElectronsEq") == dump()
    o0 = ctx.observe()[0]
    ctx.step(dt, steps, seed=seed)
    assert read("ElectronsEnd") == dump()
    av = ctx.averages(0)
    row0 = dict(avgE=o0[3] / o0[2], avgV=o0[4] / o0[2], occ=o0[2] / n)
    for part, key in (("AvgEnergy", "avgE"), ("AvgDriftVelocity", "avgV"),
                      ("valleyOccupation", "occ")):
        vals = [row0[key]] + list(av[key][:, 0])
        text = "".join("%g %g\n" % (s * dt, v) for s, v in enumerate(vals))
        assert read(part) == text.encode(), part


@pytest.fixture(scope="module")
def output_binary(tmp_path_factory):
    out = str(tmp_path_factory.mktemp("cpp") / "output_format")
    subprocess.check_call(["g++", "-std=c++17", "-O2", "-I", os.path.join(PKG, "include"),
                           os.path.join(ROOT, "tests", "cpp", "output_format.cpp"), "-o", out])
    return out


@pytest.mark.parametrize("name,mode", [
    ("sim_PotentialEq", "grid"), ("sim_PotentialAvg", "grid"), ("sim_EFieldXEq", "grid"),
    ("sim_EFieldYEq", "grid"), ("sim_ElectronsConcEq", "grid"), ("sim_ElectronsConcAvg", "grid"),
    ("sim_ElectronsCurrent", "current"), ("sim_ElectronsEq", "particles"),
    ("sim_ElectronsFinal", "particles")])
def test_result_files_match_the_reference_format(output_binary, golden, tmp_path, name, mode):
    text = bytes(golden["simfile." + name])
    src = tmp_path / (name + ".txt")
    src.write_bytes(text)
    args = [output_binary, mode, str(src)]
    if mode == "current":
        _, t_sim, dt, t_tr, _, _ = golden["sim.cfg"]
        args += ["2", repr(float(dt)), repr(float(t_tr))]
    assert subprocess.check_output(args) == text


BULK_FILES = ["si_jacoboni", "si_vasileska", "polar_gamma", "two_valley", "para_iso",
              "screened_polar"]


@pytest.mark.parametrize("name", BULK_FILES)
@pytest.mark.parametrize("part,mode", [
    ("ElectronsEq", "bulkparticles"), ("ElectronsEnd", "bulkparticles"),
    ("AvgEnergy", "timeseries"), ("AvgDriftVelocity", "timeseries"),
    ("valleyOccupation", "timeseries")])
def test_bulk_files_match_the_reference_format(output_binary, golden, tmp_path, name, part, mode):
    """The bulk example's files (print of the bulk handler and the three
    time series of bulkSimulation.cpp) re-printed through the host API's
    writers equal the reference's files byte for byte."""
    text = bytes(golden["bulkfile.%s.%s" % (name, part)])
    src = tmp_path / (part + ".txt")
    src.write_bytes(text)
    args = [output_binary, mode, str(src)]
    if mode == "timeseries":
        args.append(repr(float(golden[name + ".handler.geom"][0])))
    assert subprocess.check_output(args) == text


def test_screened_mechanisms_match_reference(tmp_path, golden):
    """emcScreenedFroehlich{Absorption,Emission}3D and emcPlasmonScreening in
    the host API: rates bit-exact against the reference's screened_polar
    tables, q_s^2 handed to the C ABI, the Debye update and the disabled
    screening of emcPlasmonScreening.hpp:69-82."""
    exe = str(tmp_path / "host_api_screened")
    subprocess.check_call(["g++", "-std=c++17", "-O2", "-I", os.path.join(PKG, "include"),
                           os.path.join(ROOT, "tests", "cpp", "host_api_screened.cpp"), "-o", exe,
                           "-L", PKG, "-lemcgpu", "-Wl,-rpath," + PKG])
    txt = subprocess.check_output([exe]).decode().splitlines()
    rates = np.array([float(l.split()[1]) for l in txt if l.startswith("rate ")])
    ref = golden["screened_polar.rate_table"]
    assert rates.shape == ref.shape and (rates == ref).all()
    val = {l.split()[0]: l.split()[1:] for l in txt if not l.startswith("rate ")}
    assert float(val["ctx_qs2"][0]) == 1e17
    assert [int(x) for x in val["kinds"]] == [9, 10]
    Q, EPS0, KB = 1.60219e-19, 8.85419e-12, 1.38066e-23
    qs2 = 1e24 * Q * Q / (10.0 * EPS0 * KB * 600.0)
    assert abs(float(val["debye"][0]) - qs2) <= 1e-15 * qs2
    assert abs(float(val["debye"][1]) - qs2 ** -0.5) <= 1e-15 * qs2 ** -0.5
    assert float(val["off"][0]) == 0.0


@pytest.mark.gpu
@pytest.mark.parametrize("variant,name", [("plain", "hot_polar"), ("screened", "screened_hot_polar")])
def test_cpp_hot_phonon_example_matches_python_binding(variant, name):
    """hotPhononGpu (the hot-carrier drivers' loop written against the host
    API: emcPhononBath, emcHotPhononFroehlich* / emcScreenedHotPhononFroehlich*,
    emcPlasmonScreening) equals the Python binding on the same seed: mean
    energy and mean N_q after every step, and the final q_s^2."""
    from viennaemc_amd import emcgpu
    exe = os.path.join(PKG, "examples", "hotPhononGpu")
    n, steps, seed = 20000, 60, 97
    print("cpp hot-phonon example", variant, flush=True)
    out = subprocess.check_output([exe, str(n), str(steps), "fp64", str(seed), variant],
                                  env=dict(os.environ, EMCGPU_THROW="1"), timeout=120).decode()
    rows = [l.split() for l in out.splitlines() if l.startswith("step ")]
    assert len(rows) == steps
    spec = dict(MODELS[name])
    if variant == "screened":
        # the example's emcPlasmonScreening starts at q_s^2 = 0, and the initial
        # free flights follow the tables of that moment
        spec["screening"] = dict(spec["screening"], qs2=0.0)
    ctx = emcgpu.Context(precision=emcgpu.FP64)
    ctx.setup(0, spec)
    ctx.resize([n])
    ctx.init_particles(0, seed, spec["T"])
    if variant == "screened":
        ctx.update_screening(0)
    ctx.step(1e-15, steps, seed=seed)
    av = ctx.averages(0)
    hist = ctx.bath_history(0, 0)
    for r in rows:
        s = int(r[1]) - 1
        assert float(r[2]) == av["avgE"][s, 0]
        assert float(r[3]) == hist[s]
    if variant == "screened":
        assert float(rows[-1][4]) == ctx.get_screening(0) and ctx.get_screening(0) > 0
    assert hist[-1] != hist[0]
