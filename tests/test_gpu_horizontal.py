"""GPU horizontal step (gitm_horiz_advance) against the oracle's restatement
(oh_advance_horizontal, itself bit for bit with the reference Fortran
build): the same IEEE operations in the same order, so bit for bit on every
field the step writes (sin(latitude) is the device's; the Coriolis cases
have passed bit for bit too, and are held to 1e-13 relative otherwise);
halo cells and ghost levels are left alone; blocks of other shapes, several
altitudes per workgroup, ion advection, the golden fixture, the
negative-density report and the unsupported flags; with UseImprovedIonAdvection
the ion-velocity divergence of the second kernel (GITM's default for ions);
with UseTopography the altitude-derivative term, the levels advanced in
order (one launch each), through the device and the host entries."""
from __future__ import annotations

import os

import numpy as np
import pytest

from gitm_amd import synth
from oracle import cpu

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))
OMEGA = 2.0 * np.pi / 86400.0
STATE = ("rho", "temp", "vel", "vvel", "nds")


def run_gpu(solver, g, dt, coriolis=True, ions=False, apw=0, temp_source=True,
            improved=False, topography=False):
    import torch
    d = {k: torch.from_numpy(np.ascontiguousarray(v)).cuda() for k, v in g.items()}
    if temp_source:
        d["temp_source"] = torch.zeros_like(d["rho"])
    solver.horizontal(d, dt, OMEGA, 2.0, use_coriolis=coriolis, use_ion_advection=ions,
                      altitudes_per_workgroup=apw, use_improved_ion_advection=improved,
                      use_topography=topography)
    torch.cuda.synchronize()
    solver.check()
    return {k: v.cpu().numpy() for k, v in d.items()}


def compare(out, ref, keys, coriolis):
    for k in keys:
        a, b = out[k], ref[k]
        if k == "temp_source":
            a, b = a[:, 2:-2, 2:-2, 2:-2], b[:, 2:-2, 2:-2, 2:-2]
        if np.array_equal(a, b):
            continue
        assert coriolis, f"{k} differs without Coriolis"
        np.testing.assert_allclose(a, b, rtol=1e-13, atol=0, err_msg=k)


@pytest.mark.parametrize("spec,coriolis,ions,apw,improved", [
    (synth.BlockGrid(4, 2, 50), True, False, 0, False),
    (synth.BlockGrid(4, 2, 50), False, True, 0, False),
    (synth.BlockGrid(4, 2, 50), True, True, 0, True),
    (synth.BlockGrid(3, 2, 30, nlons=6, nlats=5), True, True, 3, True),
    (synth.BlockGrid(2, 2, 20, nlons=16, nlats=16), True, False, 1, False),
    (synth.BlockGrid(2, 1, 7, nlons=4, nlats=9), False, False, 4, False),
])
def test_matches_oracle(solver, spec, coriolis, ions, apw, improved):
    g = synth.block_grid(spec, seed=21, ions=ions)
    dt = synth.horizontal_dt(g, cfl=0.4)
    print(f"horizontal {spec} coriolis={coriolis} ions={ions} apw={apw} improved={improved}",
          flush=True)
    out = run_gpu(solver, g, dt, coriolis, ions, apw, improved=improved)
    ref, rc, _ = cpu.horizontal(g, dt, OMEGA, 2.0, coriolis, ions, improved)
    assert rc == 0
    keys = STATE + ("temp_source",) + (("ions",) if ions else ())
    compare(out, ref, keys, coriolis)
    for k in STATE:  # halo cells and ghost levels untouched
        mask = np.ones(g[k].shape[-3:], bool)
        mask[2:-2, 2:-2, 2:-2] = False
        assert np.array_equal(out[k][..., mask], g[k][..., mask]), k


@pytest.mark.parametrize("spec,ions,apw", [
    (synth.BlockGrid(4, 2, 30), False, 0),
    (synth.BlockGrid(4, 2, 30), True, 0),
    (synth.BlockGrid(3, 2, 12, nlons=6, nlats=5), True, 3),  # apw is 1 with topography
    (synth.BlockGrid(2, 1, 8), False, 9),  # 9 x 81 lanes would not fit; ignored too
])
def test_topography_matches_oracle(solver, spec, ions, apw):
    g = synth.block_grid(spec, seed=23 + int(ions), ions=ions, topography=True)
    dt = synth.horizontal_dt(g, cfl=0.4)
    print(f"horizontal topography {spec} ions={ions} apw={apw}", flush=True)
    out = run_gpu(solver, g, dt, True, ions, apw, improved=ions, topography=True)
    ref, rc, _ = cpu.horizontal(g, dt, OMEGA, 2.0, True, ions, ions, use_topography=True)
    assert rc == 0
    keys = STATE + ("temp_source",) + (("ions",) if ions else ())
    compare(out, ref, keys, True)
    for k in STATE:
        mask = np.ones(g[k].shape[-3:], bool)
        mask[2:-2, 2:-2, 2:-2] = False
        assert np.array_equal(out[k][..., mask], g[k][..., mask]), k


def test_golden_fixture(solver):
    import importlib.util
    spec = importlib.util.spec_from_file_location(
        "mkh", os.path.join(HERE, "golden", "make_golden_horizontal.py"))
    mk = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mk)
    fx = np.load(os.path.join(HERE, "golden", mk.NAME))
    g, dt = mk.inputs()
    assert str(fx["inputs_sha256"]) == mk.digest(g)
    out = run_gpu(solver, g, dt, True, True, improved=True)
    lv = fx["levels"]
    for k in STATE + ("ions",):
        got = out[k][..., lv + 1, 2:-2, 2:-2]
        if not np.array_equal(got, fx["out_" + k]):
            np.testing.assert_allclose(got, fx["out_" + k], rtol=1e-13, atol=0, err_msg=k)


def test_one_degree_sample(solver):
    """Six blocks of the 1-degree grid (100 levels; polar, mid-latitude and
    equatorial), against the oracle."""
    full = synth.BLOCK_CONFIGS["1deg_blocks_800x100"]
    g = synth.block_grid(full, seed=3, blocks=[0, 1, 39, 200, 410, 799])
    dt = synth.horizontal_dt(g, cfl=0.4)
    out = run_gpu(solver, g, dt)
    ref, rc, _ = cpu.horizontal(g, dt)
    assert rc == 0 and full.ncells == 64800
    compare(out, ref, STATE + ("temp_source",), True)


def test_negative_density_is_reported(solver):
    import torch
    from gitm_amd.vertical import VerticalStepError
    g = synth.block_grid(synth.BlockGrid(1, 1, 10), seed=2)
    g["nds"][0, 3, 5, 6, 4] = -1.0e12
    d = {k: torch.from_numpy(v).cuda() for k, v in g.items()}
    solver.horizontal(d, 1.0)
    torch.cuda.synchronize()
    with pytest.raises(VerticalStepError) as e:
        solver.check()
    assert e.value.level == 4 and e.value.column == (6 - 2) * 9 + (4 - 2)


def test_unsupported_flags(solver):
    import torch
    from gitm_amd import vertical
    g = synth.block_grid(synth.BlockGrid(1, 1, 10), seed=2, ions=True)
    d = {k: torch.from_numpy(v).cuda() for k, v in g.items()}
    with pytest.raises(vertical.VerticalStepError) as e:
        solver.horizontal(d, 1.0, use_ion_advection=True, altitudes_per_workgroup=9)
    assert e.value.code == 4  # 9 x 81 lanes do not fit a workgroup
    del d["grad_lat"]
    with pytest.raises(KeyError):  # the divergence needs UAM_Gradient's coefficients
        solver.horizontal(d, 1.0, use_ion_advection=True, use_improved_ion_advection=True)
    c, f, p = vertical._horiz_structs({k: t.data_ptr() for k, t in d.items()},
                                      tuple(d["rho"].shape), 1.0, OMEGA, 2.0, True, False, 0,
                                      False)
    p.use_topography = 1  # without dAlt_GB and the slopes
    assert vertical.lib().gitm_horiz_advance(solver._h, p, c, f) == 4


def test_host_entry_matches_device_entry(solver):
    """gitm_horiz_advance_host (host arrays, staged over the ranges the step
    touches) leaves the same arrays as the device entry, ions and the
    improved ion advection included; the inputs are left alone."""
    g = synth.block_grid(synth.BlockGrid(3, 2, 20), seed=31, ions=True)
    dt = synth.horizontal_dt(g, cfl=0.4)
    dev = run_gpu(solver, g, dt, True, True, improved=True)  # adds a zero temp_source
    host = {k: np.array(v) for k, v in g.items()}
    host["temp_source"] = np.zeros_like(g["rho"])
    solver.horizontal_host(host, dt, use_coriolis=True, use_ion_advection=True,
                           use_improved_ion_advection=True)
    for k in ("rho", "temp", "vel", "vvel", "nds", "ions"):
        assert np.array_equal(host[k], dev[k]), k
    assert np.array_equal(host["temp_source"], dev["temp_source"])
    for k in ("gamma", "cmax", "raddist", "ivel_par"):
        assert np.array_equal(host[k], g[k]), k


def test_host_entry_topography(solver):
    """The host entry with UseTopography stages dAlt_GB and the slopes and
    gives the device entry's arrays."""
    g = synth.block_grid(synth.BlockGrid(2, 2, 16), seed=37, ions=True, topography=True)
    dt = synth.horizontal_dt(g, cfl=0.4)
    dev = run_gpu(solver, g, dt, True, True, improved=True, topography=True)
    host = {k: np.array(v) for k, v in g.items()}
    host["temp_source"] = np.zeros_like(g["rho"])
    solver.horizontal_host(host, dt, use_coriolis=True, use_ion_advection=True,
                           use_improved_ion_advection=True, use_topography=True)
    for k in ("rho", "temp", "vel", "vvel", "nds", "ions", "temp_source"):
        assert np.array_equal(host[k], dev[k]), k
    for k in ("dalt", "dalt_dlon", "dalt_dlat"):
        assert np.array_equal(host[k], g[k]), k
