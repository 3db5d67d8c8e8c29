"""GPU column prep and vertical time step (gitm_vert_prepare) against the
oracle's restatement (ov_prepare), and the prep chained into the step."""
from __future__ import annotations

import dataclasses

import numpy as np
import pytest

from gitm_amd import synth
from oracle import cpu
from tests.helpers import assert_parity, subset

pytestmark = pytest.mark.gpu

PREP = ("nden", "press", "cp", "gamma", "mmm", "tempunit", "eddy")
# sums, products, quotients and sqrt in the reference's order: the device's
# IEEE operations give the oracle's bits; 1 ulp is allowed for sqrt
RTOL = 2.3e-16


def gpu_prepare(solver, cols, pp):
    import torch
    from gitm_amd import vertical
    d = vertical.to_device(cols)
    L = d["temp"].shape[1]
    for k in PREP:
        d[k] = torch.zeros_like(d["temp"])
    d["cmax"] = torch.zeros_like(d["temp"])
    dt = solver.prepare(d, pp)
    out = vertical.to_host(d)
    return out, float(dt.item())


def rel(a, b):
    return float(np.max(np.abs(a - b) / np.maximum(np.abs(b), 1e-300)))


@pytest.mark.parametrize("config,ions", [("cpu_324x50", False), ("cpu_324x50", True),
                                         ("1deg_64800x100", False)])
def test_prepare_matches_oracle(solver, config, ions):
    spec = synth.CONFIGS[config]
    cols = synth.make_columns(spec, seed=21, ions=ions)
    if spec.ncols > 4096:
        cols = subset(cols, np.random.default_rng(3).choice(spec.ncols, 512, replace=False))
    pp = dataclasses.replace(synth.PrepParams(), fixed_dt=1e9, use_ion_advection=ions)
    out, dt = gpu_prepare(solver, cols, pp)
    cols0 = dict(cols)
    cols0["cmax"] = np.zeros_like(cols["temp"])
    ref, dt_ref = cpu.prepare(cols0, pp)
    errs = {k: rel(out[k], ref[k]) for k in PREP}
    errs["cmax"] = rel(out["cmax"][:, 1:-1], ref["cmax"][:, 1:-1])
    bad = {k: v for k, v in errs.items() if not v <= RTOL}
    assert not bad, (bad, errs)
    assert abs(dt - dt_ref) <= RTOL * dt_ref, (dt, dt_ref)
    exact = [k for k in PREP if np.array_equal(out[k], ref[k])]
    print(config, "bitwise:", exact, "max rel:", errs, "dt", dt, dt_ref)


def test_prepare_caps(solver):
    cols = subset(synth.make_columns(synth.CONFIGS["cpu_324x50"], seed=22), np.arange(40))
    pp = synth.PrepParams()
    _, dt = gpu_prepare(solver, cols, dataclasses.replace(pp, fixed_dt=1e-3))
    assert dt == 1e-3
    _, dt = gpu_prepare(solver, cols, dataclasses.replace(pp, fixed_dt=1e9, dt_end=2e-3))
    assert dt == 2e-3


def test_prepare_then_step(solver):
    """Prep on the device, then one AUSM step from its outputs, against the
    oracle doing the same (the step reads TempUnit, MeanMajorMass, Gamma,
    cMax and the eddy coefficient from the prep)."""
    from gitm_amd import vertical
    import torch
    cols = subset(synth.make_columns(synth.CONFIGS["cpu_324x50"], seed=23), np.arange(0, 324, 4))
    pp = synth.PrepParams()
    d = vertical.to_device(cols)
    d["press"] = torch.zeros_like(d["temp"])
    d["cp"] = torch.zeros_like(d["temp"])
    solver.prepare(d, pp)
    solver.advance(d, synth.StepParams(), "ausm", 1)
    out = vertical.to_host(d)
    ref_prep, _ = cpu.prepare(cols, pp)
    chained = dict(cols)
    for k in ("gamma", "mmm", "tempunit", "eddy"):
        chained[k] = ref_prep[k]
    chained["cmax"] = ref_prep["cmax"]
    ref = cpu.advance(chained, synth.StepParams(), solver=1, threads=8)
    assert_parity(out, ref, 1e-11, what="prep + step")
