"""In-tree build of the HIP library libemcgpu.so for gfx950.

Compiles the kernels and the C ABI layer with hipcc into
viennaemc_amd/libemcgpu.so.  Runs on a machine without a GPU (hipcc
cross-compiles).  Packed-FP32 instruction selection is disabled for all
device code.
"""
import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "libemcgpu.so")

SOURCES = ["emcgpu_kernels.hip", "emcgpu_capi.cpp", "emcgpu_device_kernels.hip",
           "emcgpu_device_capi.cpp"]
HEADERS = ["emcgpu_device.h", "emcgpu_particle.h", "emcgpu_host.h", "../include/emcgpu/emcgpuPhysics.hpp", "../../include/emcgpu.h"]

ARCH = os.environ.get("EMCGPU_ARCH", "gfx950")


def hipcc():
    for cand in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if cand and os.path.exists(cand):
            return cand
    raise RuntimeError("hipcc not found")


def flags():
    return [
        "--offload-arch=%s" % ARCH, "-O3", "-std=c++17", "-fPIC", "-Wall",
        "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
        "-munsafe-fp-atomics",
    ]


def build(verbose=True):
    """Compile every source and link libemcgpu.so (always recompiles)."""
    objdir = os.path.join(HERE, "_build")
    os.makedirs(objdir, exist_ok=True)
    objs = []
    cc = hipcc()
    for src in SOURCES:
        obj = os.path.join(objdir, os.path.splitext(src)[0] + ".o")
        cmd = [cc] + flags() + ["-c", os.path.join(CSRC, src), "-o", obj]
        if src.endswith(".cpp"):
            cmd = [cc] + flags() + ["-x", "hip", "-c", os.path.join(CSRC, src), "-o", obj]
        if verbose:
            print(" ".join(cmd), flush=True)
        subprocess.check_call(cmd)
        objs.append(obj)
    tmp = LIB + ".tmp"
    cmd = [cc, "--offload-arch=%s" % ARCH, "-shared", "-fPIC", "-o", tmp] + objs + [
        "-lrocsolver", "-lrocblas"]
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.check_call(cmd)
    os.replace(tmp, LIB)
    build_examples(verbose)
    build_calibration(verbose)
    return LIB


CALIB = os.path.join(HERE, "calib", "valu_ceiling")


def build_calibration(verbose=True):
    """The vector-ALU issue-rate calibration program (calib/valu_ceiling.hip),
    run by bench.py --valu-calibration."""
    cmd = [hipcc()] + [f for f in flags() if f not in ("-fPIC", "-munsafe-fp-atomics")] + [
        CALIB + ".hip", "-o", CALIB]
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.check_call(cmd)


EXAMPLES = ["bulkSimulationGpu", "resistor2DGpu", "hotPhononGpu"]


def build_examples(verbose=True):
    """Host-only C++17 programs written against the header-only host API,
    linked to libemcgpu.so (rpath to the package folder)."""
    for name in EXAMPLES:
        src = os.path.join(HERE, "examples", name + ".cpp")
        out = os.path.join(HERE, "examples", name)
        cmd = ["g++", "-std=c++17", "-O2", "-Wall", "-I", os.path.join(HERE, "include"), src,
               "-o", out, "-L", HERE, "-lemcgpu", "-Wl,-rpath,$ORIGIN/.."]
        if verbose:
            print(" ".join(cmd), flush=True)
        subprocess.check_call(cmd)


if __name__ == "__main__":
    build()
    sys.exit(0)
