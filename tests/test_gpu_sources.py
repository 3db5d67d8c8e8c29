"""GPU sources step (gitm_vert_sources) against the oracle's restatement
(ov_sources_step): the same IEEE additions, multiplications and divisions
in the same order, so bit for bit; increments and the updated state, with
and without turbulent conduction, on uniform and stretched grids."""
from __future__ import annotations

import numpy as np
import pytest

from gitm_amd import synth
from oracle import cpu
from tests.helpers import subset

pytestmark = pytest.mark.gpu

STATE = ("temp", "vel", "vvel")
INC = ("conduction", "viscosity", "vert_viscosity")


def gpu_sources(solver, cols, src, dt, apply, turbulent):
    import torch
    from gitm_amd import vertical
    d = {k: torch.from_numpy(np.ascontiguousarray(cols[k])).cuda()
         for k in ("temp", "tempunit", "rho", "eddy", "nds", "vel", "vvel", "alt", "invr",
                   "grav", "dalt", "mmm", "gamma", "cmax", "coslat")}
    s = {k: torch.from_numpy(np.ascontiguousarray(v)).cuda() for k, v in src.items()}
    s["conduction"] = torch.zeros_like(d["temp"])
    s["viscosity"] = torch.zeros_like(d["vel"])
    s["vert_viscosity"] = torch.zeros_like(d["vvel"])
    solver.sources(d, s, dt, apply=apply, use_turbulent_cond=turbulent)
    torch.cuda.synchronize()
    out = {k: d[k].cpu().numpy() for k in STATE}
    out.update({k: s[k].cpu().numpy() for k in INC})
    return out


@pytest.mark.parametrize("config,apply,turbulent", [
    ("cpu_324x50", True, False), ("cpu_324x50", False, True), ("1deg_64800x100", True, True),
    ("short", True, False), ("max", True, True)])
def test_sources_match_oracle(solver, config, apply, turbulent):
    # "short": 10 levels, fewer lanes than the 18 concurrent solves; "max":
    # the largest column (252 levels, four wavefronts, stretched grid)
    small = {"short": synth.GridSpec(1, 1, 10),
             "max": synth.GridSpec(1, 1, 252, nlons=3, nlats=2, stretched=True)}
    spec = small[config] if config in small else synth.CONFIGS[config]
    cols = synth.make_columns(spec, seed=51)
    if spec.ncols > 4096:
        cols = subset(cols, np.random.default_rng(5).choice(spec.ncols, 300, replace=False))
    src = synth.source_fields(cols, seed=51)
    mass = synth.species_masses()
    out = gpu_sources(solver, cols, src, 2.0, apply, turbulent)
    ref = cpu.sources(dict(cols, **src), 2.0, mass, use_turbulent_cond=turbulent, apply=apply)
    L = cols["temp"].shape[1]
    for k in STATE:
        assert np.array_equal(out[k], ref[k]), k
    for k in INC:  # levels 0..nAlts+1
        assert np.array_equal(out[k][..., 1:L - 1], ref[k][..., 1:L - 1]), k
    assert np.max(np.abs(out["conduction"])) > 0 and np.max(np.abs(out["vert_viscosity"])) > 0


def test_host_entry_matches_device_entry(solver):
    """gitm_vert_sources_host (host arrays, staged over the levels the step
    touches) leaves the same state and increments as the device entry."""
    cols = synth.make_columns(synth.CONFIGS["cpu_324x50"], seed=52)
    src = synth.source_fields(cols, seed=52)
    dev = gpu_sources(solver, cols, src, 2.0, True, True)
    host = {k: np.array(cols[k]) for k in ("temp", "tempunit", "rho", "eddy", "nds", "vel",
                                           "vvel", "alt")}
    hs = {k: np.array(v) for k, v in src.items()}
    hs["conduction"] = np.zeros_like(host["temp"])
    hs["viscosity"] = np.zeros_like(host["vel"])
    hs["vert_viscosity"] = np.zeros_like(host["vvel"])
    solver.sources_host(host, hs, 2.0, apply=True, use_turbulent_cond=True)
    for k in STATE:
        assert np.array_equal(host[k], dev[k]), k
    L = host["temp"].shape[1]
    for k in INC:
        assert np.array_equal(hs[k][..., 1:L - 1], dev[k][..., 1:L - 1]), k
