"""The production-size workloads at their full sizes on the GPU, each step
checked against the oracle on a sample of columns (columns never interact,
so a sample is checked exactly; the rest is checked for finiteness,
positivity and bitwise reproducibility).

* 1 deg x 1 deg x 100 stretched levels, 64,800 columns, 50 AUSM steps (one
  launch per step): before every step the sampled columns are copied off the
  device, the oracle takes the same step from them, and the GPU's step must
  agree at 1e-11; 50 steps in one launch must equal the 50 launches bit for
  bit.
* The 64-member ragged ensemble, 725,760 columns, 20 x (batch prep with
  per-member CFL steps + batch step): per step, two columns of every nAlts
  group (with the prep outputs and the member's step the GPU used) against
  the oracle at 1e-11; the per-member CFL steps of the initial state (with
  FixedDt lifted, so that they differ) against the oracle's prep; then the
  same 20 steps as one hipGraph replayed from the initial state must equal
  the eager loop bit for bit for every member.
"""
from __future__ import annotations

import os

import numpy as np
import pytest

from gitm_amd import synth
from oracle import cpu
from tests.helpers import assert_parity

pytestmark = pytest.mark.gpu

RTOL = 1e-11
THREADS = min(16, os.cpu_count() or 1)
STATE = ("nds", "vvel", "vel", "temp", "rho")


def take(d: dict, idx) -> dict:
    """Host copies of the columns ``idx`` of the device columns ``d``."""
    return {k: v[idx].cpu().numpy() for k, v in d.items()}


def test_1deg_grid_50_steps_sampled_per_step(solver):
    import torch
    from gitm_amd import vertical
    spec = synth.CONFIGS["1deg_64800x100"]
    cols = synth.make_columns(spec, seed=11)
    p = synth.StepParams()
    d = vertical.to_device(cols)
    init = {k: v.clone() for k, v in d.items()}
    rng = np.random.default_rng(50)
    idx = torch.from_numpy(np.sort(rng.choice(spec.ncols, 48, replace=False))).cuda()
    worst = {}
    for step in range(50):
        before = take(d, idx)
        solver.advance(d, p, "ausm", 1)
        after = take(d, idx)
        ref = cpu.advance(before, p, solver=1, threads=THREADS)
        errs = assert_parity(after, ref, RTOL, what=f"1deg step {step + 1}")
        for k, v in errs.items():
            worst[k] = max(worst.get(k, 0.0), v)
    print("1deg 50 steps, worst per-step error", worst, flush=True)
    for k in ("nds", "temp", "rho", "nden"):
        assert bool((d[k] > 0).all()) and bool(torch.isfinite(d[k]).all()), k
    for k in ("vvel", "vel", "logns"):
        assert bool(torch.isfinite(d[k]).all()), k
    solver.advance(init, p, "ausm", 50)
    for k in STATE:
        assert torch.equal(init[k], d[k]), f"50 steps in one launch differ in {k}"


def ensemble_members(torch, seed=4):
    """The 64 members on the device as bench.py builds them: each member's
    columns tile the first min(n, 2592) columns of its own seeded
    atmosphere.  Also returns those host columns."""
    from gitm_amd import vertical
    mem = synth.ensemble(64, seed=seed)
    dev, host = [], []
    for m in mem:
        n = m.spec.ncols
        h = synth.member_columns(m, min(n, 2592))
        t = vertical.to_device(h)
        reps = -(-n // t["temp"].shape[0])
        dev.append({k: v.repeat((reps,) + (1,) * (v.dim() - 1))[:n].contiguous()
                    for k, v in t.items()})
        host.append(h)
        del t
    return dev, host


def test_full_ensemble_20_steps(solver):
    import dataclasses

    import torch
    from gitm_amd import vertical
    dev, host = ensemble_members(torch)
    ncols = sum(int(d["temp"].shape[0]) for d in dev)
    assert len(dev) == 64 and ncols == 725760
    init = [{k: v.clone() for k, v in d.items()} for d in dev]
    p = synth.StepParams()
    pp = synth.PrepParams()
    b = vertical.VerticalBatch(solver, dev)
    print("ensemble tile", b.tile(), flush=True)
    # two columns of one member of every nAlts group
    groups = {}
    for i, d in enumerate(dev):
        groups.setdefault(int(d["temp"].shape[1]) - 4, i)
    assert sorted(groups) == [50, 60, 70, 80, 90, 100, 110, 120]
    rng = np.random.default_rng(20)
    sample = {i: torch.from_numpy(np.sort(rng.choice(int(dev[i]["temp"].shape[0]), 2,
                                                     replace=False))).cuda()
              for i in groups.values()}
    dt = torch.empty(len(dev), dtype=torch.float64, device="cuda")
    # each member's CFL step: the oracle's prep over its distinct columns
    pp_cfl = dataclasses.replace(pp, fixed_dt=1e9)
    got = b.prepare(pp_cfl, dt_out=dt).cpu().numpy()
    want = np.array([cpu.prepare(h, pp_cfl)[1] for h in host])
    assert np.array_equal(got, want), (got, want)
    assert len(set(got.tolist())) > 1
    for step in range(20):
        b.prepare(pp, dt_out=dt)
        dts = dt.cpu().numpy()
        before = {i: take(dev[i], idx) for i, idx in sample.items()}
        b.advance(p, "ausm", 1, dt=dt)
        for i, idx in sample.items():
            after = take(dev[i], idx)
            pi = dataclasses.replace(p, dt=float(dts[i]))
            ref = cpu.advance(before[i], pi, solver=1, threads=THREADS)
            assert_parity(after, ref, RTOL,
                          what=f"member {i} (nAlts {dev[i]['temp'].shape[1] - 4}) "
                               f"step {step + 1}")
    for d in dev:
        assert bool(torch.isfinite(d["temp"]).all()) and bool((d["nds"] > 0).all())
    eager = [{k: d[k].clone() for k in STATE} for d in dev]
    # the same 20 steps as one graph, replayed from the initial state
    for d, d0 in zip(dev, init):
        for k in d:
            d[k].copy_(d0[k])
    torch.cuda.synchronize()
    g = torch.cuda.CUDAGraph()
    with torch.cuda.graph(g):
        for _ in range(20):
            b.prepare(pp, dt_out=dt)
            b.advance(p, "ausm", 1, dt=dt, check=False)
    for d, d0 in zip(dev, init):  # capture queues nothing; start from the initial state
        for k in d:
            d[k].copy_(d0[k])
    torch.cuda.synchronize()
    g.replay()
    torch.cuda.synchronize()
    b.check()
    for i, (d, e) in enumerate(zip(dev, eager)):
        for k in STATE:
            assert torch.equal(d[k], e[k]), f"graph replay differs: member {i}, {k}"
    del g
    b.close()
