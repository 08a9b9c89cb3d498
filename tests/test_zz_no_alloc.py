"""emcgpu_step_noalloc against emcgpu_step, through the C ABI.

The allocation-free step keeps its per-step partials in the library's static
scratch and writes the observables to a device buffer of the caller.  Two
calls in a row on the shared scratch (the second with fewer particles, so
the scratch still holds the first call's larger contents), each against
emcgpu_step on an identical context: final particle state bit for bit and
every per-step sum exactly.  The first call asks for more steps per launch
than the scratch holds partials for, so the cap on the launch shape is
exercised as well.
"""
import ctypes as C

import numpy as np
import pytest

from emc_models import MODELS

pytestmark = pytest.mark.gpu

DT = 1e-15


def _hip():
    """The HIP runtime libemcgpu.so is linked against (the copy already
    mapped into this process), for the caller-side device buffer."""
    from viennaemc_amd import emcgpu
    emcgpu.load()
    path = None
    with open("/proc/self/maps") as f:
        for line in f:
            if "libamdhip64.so" in line:
                path = line.split()[-1]
                break
    assert path, "libamdhip64 is not mapped"
    H = C.CDLL(path)
    H.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
    H.hipFree.argtypes = [C.c_void_p]
    H.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
    H.hipMemset.argtypes = [C.c_void_p, C.c_int, C.c_size_t]
    for f in (H.hipMalloc, H.hipFree, H.hipMemcpy, H.hipMemset):
        f.restype = C.c_int
    return H


HIP_D2H = 2


def _ctx(emcgpu, precision, n, seed):
    spec = MODELS["si_jacoboni"]
    ctx = emcgpu.Context(precision=precision)
    ctx.setup(0, spec)
    ctx.resize([n])
    ctx.init_particles(0, seed, spec["T"])  # puts the model on the device at 1 fs
    return ctx


@pytest.mark.parametrize("precision", [0, 1])
def test_noalloc_step_matches_step(precision):
    from viennaemc_amd import emcgpu
    H = _hip()
    # (particles, steps, steps per launch asked): 4M particles need 1 MiB of
    # partials per step, so 128 steps per launch are capped to the scratch
    calls = [(4_000_000, 150, 128), (100_003, 40, 0)]
    for n, steps, spl in calls:
        seed = 1000 + n
        print("noalloc vs step: precision %d, %d particles, %d steps" % (precision, n, steps),
              flush=True)
        ref = _ctx(emcgpu, precision, n, seed)
        ref.step(DT, steps, seed=seed, step0=3, steps_per_launch=spl)
        want = ref.get_particles(0)
        want_obs = ref.observables(0)
        ref.close()

        ctx = _ctx(emcgpu, precision, n, seed)
        n_obs = ctx.obs_layout()[0]
        m = steps * n_obs
        obs = C.c_void_p()
        assert H.hipMalloc(C.byref(obs), m * 8) == 0
        try:
            assert H.hipMemset(obs, 0xFF, m * 8) == 0  # NaN until written
            # a buffer one value short is refused before any work
            with pytest.raises(emcgpu.EmcGpuError):
                ctx.step_noalloc(DT, steps, obs.value, m - 1, seed=seed, step0=3,
                                 steps_per_launch=spl)
            ctx.step_noalloc(DT, steps, obs.value, m, seed=seed, step0=3, steps_per_launch=spl)
            ctx.synchronize()
            got = ctx.get_particles(0)
            got_obs = np.zeros(m)
            assert H.hipMemcpy(got_obs.ctypes.data_as(C.c_void_p), obs, m * 8, HIP_D2H) == 0
            got_obs = got_obs.reshape(steps, n_obs)
        finally:
            H.hipFree(obs)
        ctx.close()
        for key in ("k", "E", "tau", "valley", "sub"):
            assert np.array_equal(got[key], want[key]), key
        assert np.array_equal(got_obs, want_obs)
