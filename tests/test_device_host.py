"""Device geometry of the C ABI on a host-only context (no GPU).

emcgpu_device_setup builds the emcDevice geometry on the host: extent,
doping regions applied in order, ohmic contacts marked on every boundary
their nodes lie on, thermal voltage, Debye length and cell volume.  Checked
node by node against the oracle's restatement (itself pinned to the
reference's device fixtures), on the golden 21 x 7 resistor, the 200 x 50
workload and a device with contacts on three boundaries meeting at
corners.  The C structures of include/emcgpu.h are laid out as the binding
declares them (sizes and offsets printed by the C compiler).
"""
import ctypes as C
import os
import subprocess
import tempfile

import numpy as np
import pytest

from emc_models import (DEVICE_SMALL, MODELS_DEVICE, SI_EPS_R, SI_NI, device_region_specs,
                        mosfet_device, resistor_device)
from oracle import oracle_py as O
from viennaemc_amd import emcgpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))

CORNERS = dict(
    max_pos=(12e-9, 8e-9), spacing=(1e-9, 0.5e-9), width=2e-6, T=250.0, eps_r=SI_EPS_R, ni=SI_NI,
    doping=[((0, 0), (12e-9, 8e-9), 3e23), ((2e-9, 1e-9), (7e-9, 5e-9), 2e25),
            ((10e-9, 0), (14e-9, 3e-9), 1e21)],
    contacts=[(2, 0.1, 0.0, 4e-9), (0, -0.3, 2e-9, 8e-9), (3, 0.25, 6e-9, 12e-9)])


def _host(dev):
    ctx = emcgpu.Context(precision=emcgpu.FP64, device=-1)
    ctx.setup(0, MODELS_DEVICE["device_para"])
    ctx.device_setup(dev)
    return ctx


@pytest.mark.parametrize("dev", [DEVICE_SMALL, resistor_device(), CORNERS],
                         ids=["small", "resistor", "corners"])
def test_geometry_matches_oracle(dev):
    ctx = _host(dev)
    od = O.OracleDevice(dev)
    assert ctx.ext == tuple(od.ext)
    g, og = ctx.device_grids(), od.grids()
    assert (g["doping"] == og["doping"]).all()
    assert (g["region"] == og["region"]).all()
    assert (g["contact"] == og["contact"]).all()
    assert ctx.Vt == od.Vt and ctx.debye == od.debye
    assert ctx.cell_volume == od.cell_volume
    ctx.close()


def test_corner_contacts_cover_both_boundaries():
    """updateAllOccurences: a contact node on a corner is marked on both
    boundaries it lies on, and a node reports the contact of the first
    boundary (XMIN, XMAX, YMIN, YMAX) it lies on."""
    ctx = _host(CORNERS)
    nx, ny = ctx.ext
    con = ctx.device_grids()["contact"].reshape(ny, nx)
    assert con[0, 0] == 0 and con[0, 4] == 0       # YMIN contact, corner included
    assert (con[4:ny, 0] == 1).all()               # XMIN from y = 2 nm
    assert con[ny - 1, 0] == 1                     # XMIN reaches the top corner
    assert con[ny - 1, nx - 1] == 2                # YMAX's corner, seen from XMAX
    assert (con[1:-1, 1:-1] == -1).all()
    ctx.close()


def test_rejects_regions_and_contacts_outside():
    bad = dict(DEVICE_SMALL, doping=[((30e-9, 0), (40e-9, 6e-9), 1e24)])
    with pytest.raises(emcgpu.EmcGpuError):
        _host(bad)
    bad = dict(DEVICE_SMALL, contacts=[(0, 0.1, 0.0, 20e-9)])
    with pytest.raises(emcgpu.EmcGpuError):
        _host(bad)
    bad = dict(DEVICE_SMALL, max_pos=(1e-9, 6e-9))  # fewer than 3 nodes in x
    with pytest.raises(emcgpu.EmcGpuError):
        _host(bad)



def test_grid_limits_of_the_poisson_solver():
    """The mosfet2D example's 126 x 101 grid (multigrid levels beyond the
    LDS: kept in a global workspace) is accepted and its geometry matches
    the oracle's; a grid wider than the solver's 1024-node row map is
    rejected."""
    dev = mosfet_device()
    ctx = _host(dev)
    od = O.OracleDevice(dev)
    assert ctx.ext == (126, 101) == tuple(od.ext)
    g, og = ctx.device_grids(), od.grids()
    assert (g["doping"] == og["doping"]).all() and (g["contact"] == og["contact"]).all()
    ctx.close()
    wide = dict(DEVICE_SMALL, max_pos=(1100e-9, 3e-9),
                doping=[((0, 0), (1100e-9, 3e-9), 1e24)],
                contacts=[(0, 0.0, 0.0, 3e-9), (1, 0.0, 0.0, 3e-9)])
    with pytest.raises(emcgpu.EmcGpuError, match="1024"):
        _host(wide)


C_LAYOUT = r"""
#include <stddef.h>
#include <stdio.h>
#include "emcgpu.h"
#define P(T) printf(#T " %zu\n", sizeof(T))
#define O(T, f) printf(#T "." #f " %zu\n", offsetof(T, f))
int main(void) {
  P(emcgpu_config); P(emcgpu_valley_desc); P(emcgpu_mech_desc); P(emcgpu_bath_desc);
  P(emcgpu_step_args); P(emcgpu_doping_region); P(emcgpu_contact); P(emcgpu_device_desc);
  P(emcgpu_device_drift_args); P(emcgpu_device_step_args); P(emcgpu_autotune_args);
  O(emcgpu_mech_desc, bath); O(emcgpu_step_args, draws); O(emcgpu_device_desc, contacts);
  O(emcgpu_device_desc, poisson_tol); O(emcgpu_device_drift_args, draws);
  O(emcgpu_device_desc, init_from_potential);
  O(emcgpu_device_drift_args, tr_events); O(emcgpu_device_step_args, checksum);
  O(emcgpu_autotune_args, spl); O(emcgpu_autotune_args, res_ms_per_step);
  O(emcgpu_autotune_args, best_ms_per_step); O(emcgpu_device_step_args, poisson_interval);
  O(emcgpu_device_step_args, channel_current); O(emcgpu_device_step_args, accumulate);
  P(emcgpu_grid3); O(emcgpu_grid3, spacing); O(emcgpu_grid3, pm_scheme);
  return 0;
}
"""


def test_ctypes_layouts_match_the_c_compiler():
    with tempfile.TemporaryDirectory() as td:
        src, exe = os.path.join(td, "layout.c"), os.path.join(td, "layout")
        open(src, "w").write(C_LAYOUT)
        subprocess.check_call(["gcc", "-I", os.path.join(ROOT, "include"), src, "-o", exe])
        out = dict(line.rsplit(" ", 1) for line in subprocess.check_output([exe]).decode().split("\n")
                   if line)
    py = {"emcgpu_config": emcgpu.Config, "emcgpu_valley_desc": emcgpu.ValleyDesc,
          "emcgpu_mech_desc": emcgpu.MechDesc, "emcgpu_bath_desc": emcgpu.BathDesc,
          "emcgpu_step_args": emcgpu.StepArgs, "emcgpu_doping_region": emcgpu.DopingRegion,
          "emcgpu_contact": emcgpu.ContactDesc, "emcgpu_device_desc": emcgpu.DeviceDesc,
          "emcgpu_device_drift_args": emcgpu.DeviceDriftArgs,
          "emcgpu_device_step_args": emcgpu.DeviceStepArgs,
          "emcgpu_autotune_args": emcgpu.AutotuneArgs, "emcgpu_grid3": emcgpu.Grid3}
    for name, cls in py.items():
        assert C.sizeof(cls) == int(out[name]), name
    for key, val in out.items():
        if "." in key:
            t, f = key.split(".")
            assert getattr(py[t], f).offset == int(val), key


def test_pm3_operators_check_their_arguments():
    """The 3-D particle-mesh operators reject the 2-D-only NEC scheme,
    empty or non-positive grids, an E-field grid without interior nodes and
    a boundary status that does not fit the faces, before needing a GPU;
    valid arguments on a host-only context then report the missing GPU."""
    ctx = emcgpu.Context(precision=emcgpu.FP64, device=-1)
    ext, h = (4, 3, 3), (1e-9, 1e-9, 1e-9)
    pos = np.zeros((1, 3))
    with pytest.raises(emcgpu.EmcGpuError, match="NEC"):
        ctx.deposit3(ext, h, pos, pm=emcgpu.PM_NEC)
    with pytest.raises(emcgpu.EmcGpuError, match="ext"):
        ctx.deposit3((4, 0, 3), h, pos)
    with pytest.raises(emcgpu.EmcGpuError, match="spacing"):
        ctx.interpolate3(ext, (1e-9, -1e-9, 1e-9), *[np.zeros(36)] * 3, pos)
    bc = np.full((3, 36), -2, dtype=np.int8)
    with pytest.raises(emcgpu.EmcGpuError, match="ext"):
        ctx.efield3((4, 2, 3), h, 0.025, np.full((3, 24), -1), np.zeros(24))
    with pytest.raises(emcgpu.EmcGpuError, match="face"):
        ctx.efield3(ext, h, 0.025, bc, np.zeros(36))  # face nodes marked -2
    kk, jj, ii = np.meshgrid(range(3), range(3), range(4), indexing="ij")
    for d, c in ((0, ii), (1, jj), (2, kk)):
        bc[d, ((c == 0) | (c == ext[d] - 1)).reshape(-1)] = -1
    with pytest.raises(emcgpu.EmcGpuError, match="GPU"):
        ctx.efield3(ext, h, 0.025, bc, np.zeros(36))
    with pytest.raises(emcgpu.EmcGpuError, match="GPU"):
        ctx.deposit3(ext, h, pos)
    with pytest.raises(ValueError, match="grid"):  # host buffers of the wrong size
        ctx.deposit3(ext, h, pos, grid=np.zeros(35))
    with pytest.raises(ValueError, match="bc"):
        ctx.efield3(ext, h, 0.025, bc[:, :30], np.zeros(36))
    ctx.close()


def test_region_tables_need_an_instance_and_a_region_per_node():
    """Per-region tables: instance r holds the tables of doping region r, so
    the context needs an instance per region and every node a region."""
    specs = device_region_specs()
    ctx = emcgpu.Context(precision=emcgpu.FP64, n_instances=2, device=-1)
    for r, sp in enumerate(specs[:2]):
        ctx.setup(r, sp)
    with pytest.raises(emcgpu.EmcGpuError, match="instance per doping region"):
        ctx.device_setup(DEVICE_SMALL, region_tables=True)
    ctx.close()
    ctx = emcgpu.Context(precision=emcgpu.FP64, n_instances=3, device=-1)
    for r, sp in enumerate(specs):
        ctx.setup(r, sp)
    ctx.device_setup(DEVICE_SMALL, region_tables=True)
    holes = dict(DEVICE_SMALL, doping=DEVICE_SMALL["doping"][1:])  # the channel has no region
    with pytest.raises(emcgpu.EmcGpuError, match="lies in no doping region"):
        ctx.device_setup(holes, region_tables=True)
    ctx.close()
