"""GPU parity of the HIP vertical step against the oracle and the reference.

The oracle is the C restatement (oracle/gitm_vert_oracle.c), itself
bit-identical to the reference Fortran build on the golden fixtures
(tests/test_oracle_golden.py).  Bar: 1e-11 relative on every state variable
for every step, each GPU step starting from the oracle's state.  Densities,
temperature and rho compare elementwise; the winds, which cross zero, compare
normwise per component (max|a - b| / max|b| of each species' vertical wind and
each wind component).  The stricter elementwise ratio of the winds with a
floor at 1e-3 of the component's largest value is not a usable bar: a one-ulp
change of the inputs moves the reference's own result by ~1.3e-11 on it after
one step (measured in test_oracle_math.py::test_wind_metric_conditioning).

Free-running multi-step trajectories are held to a different bar.  The GPU's
exp/log/pow/tanh differ from glibc's in the last bit, and the AUSM step
amplifies input differences: a one-ulp change of the inputs moves the
reference's own vertical winds by ~1e-11 after one step and ~1e-9 after 100.
A trajectory therefore has to stay within twice the oracle's own divergence
under one-ulp input perturbations (``ulp_envelope``), measured in the test.
"""
from __future__ import annotations

import dataclasses
import glob
import json
import os

import numpy as np
import pytest

from gitm_amd import synth
from oracle import cpu
from tests.helpers import (assert_parity, assert_within_envelope, subset,
                           ulp_envelope)

pytestmark = pytest.mark.gpu

RTOL = 1e-11
STATE = ("nds", "vvel", "vel", "temp", "rho", "ions")
THREADS = min(16, os.cpu_count() or 1)
GOLDEN = sorted(g for g in glob.glob(os.path.join(os.path.dirname(__file__), "golden", "*.npz"))
                if not os.path.basename(g).startswith(("horizontal_", "sources_", "prep_")))
# (the vertical step's; horizontal_*, sources_* and prep_* are the other steps')


def gpu_run(solver, cols, params, which="ausm", n_steps=1):
    from gitm_amd import vertical
    d = vertical.to_device(cols)
    solver.advance(d, params, which, n_steps)
    return vertical.to_host(d)


@pytest.mark.parametrize("which", ["ausm", "rusanov"])
def test_cpu_config_one_step(solver, which):
    """324 columns (2x2 blocks of 9x9), nAlts 50, one step."""
    cols = synth.make_columns(synth.CONFIGS["cpu_324x50"], seed=1)
    p = synth.StepParams()
    out = gpu_run(solver, cols, p, which)
    ref = cpu.advance(cols, p, solver=1 if which == "ausm" else 0, threads=THREADS)
    errs = assert_parity(out, ref, RTOL, what=f"{which} 324x50")
    print(which, errs)


def per_step(solver, cols, params, which, n_steps, threads=THREADS):
    """Advance the oracle ``n_steps`` and, from the oracle's state before each
    step, one GPU step; each GPU step is held to RTOL against the oracle's.
    Returns the oracle's final state and the worst error per field."""
    from gitm_amd import vertical
    sid = 1 if which == "ausm" else 0
    state = {k: np.array(v) for k, v in cols.items()}
    worst = {}
    for step in range(n_steps):
        d = vertical.to_device(state)
        solver.advance(d, params, which, 1)
        ref = cpu.advance(state, params, solver=sid, threads=threads)
        errs = assert_parity(vertical.to_host(d), ref, RTOL,
                             what=f"{which} step {step + 1}")
        for k, v in errs.items():
            worst[k] = max(worst.get(k, 0.0), v)
        state.update({k: ref[k] for k in STATE if k in ref})
    return ref, worst


@pytest.mark.parametrize("path", GOLDEN, ids=[os.path.basename(p)[:-4] for p in GOLDEN])
def test_golden_fixture(solver, path):
    """GPU against the reference Fortran outputs stored in tests/golden: every
    step at 1e-11 (from the oracle's state, which reproduces each fixture bit
    for bit), and the free-running GPU trajectory within the reference's own
    1-ulp input sensitivity."""
    z = np.load(path)
    meta = json.loads(str(z["meta"]))
    params = synth.StepParams(**meta["params"])
    cols = {k[3:]: z[k] for k in z.files if k.startswith("in_")}
    ref = {k[4:]: z[k] for k in z.files if k.startswith("out_")}
    which = "ausm" if meta["solver"] == 1 else "rusanov"
    steps = meta["steps"]
    last, _ = per_step(solver, cols, params, which, steps)
    for k in STATE:
        if k in ref:
            assert np.array_equal(last[k], ref[k]), f"oracle differs from the fixture in {k}"
    out = gpu_run(solver, cols, params, which, steps)
    if steps == 1:
        assert_parity(out, ref, RTOL, what=os.path.basename(path))
    else:
        env = ulp_envelope(cols, params, meta["solver"], steps, ref, threads=THREADS)
        assert_within_envelope(out, ref, env, RTOL, what=os.path.basename(path))


@pytest.mark.parametrize("which", ["ausm", "rusanov"])
def test_5deg_100_steps_per_step(solver, which):
    """5 deg x 5 deg (4x8 blocks of 9x9 = 2592 columns), nAlts 50: 100
    consecutive steps, each compared with the oracle's step at 1e-11."""
    cols = synth.make_columns(synth.CONFIGS["5deg_2592x50"], seed=2)
    _, worst = per_step(solver, cols, synth.StepParams(), which, 100)
    print(which, "worst over 100 steps", worst)


@pytest.mark.parametrize("which", ["ausm", "rusanov"])
def test_5deg_100_step_trajectory(solver, which):
    """The same 100 steps run freely on the GPU (one launch of 100 steps)
    against the oracle's trajectory, on a 288-column slice: the difference
    stays within twice what a one-ulp change of the inputs does to the
    reference algorithm itself."""
    cols = subset(synth.make_columns(synth.CONFIGS["5deg_2592x50"], seed=2),
                  np.arange(0, 2592, 9))
    p = synth.StepParams()
    sid = 1 if which == "ausm" else 0
    out = gpu_run(solver, cols, p, which, 100)
    ref = cpu.advance(cols, p, solver=sid, n_steps=100, threads=THREADS)
    env = ulp_envelope(cols, p, sid, 100, ref, threads=THREADS)
    errs = assert_within_envelope(out, ref, env, RTOL, what=f"{which} 100 steps")
    print(which, "trajectory", errs, "envelope", env)


@pytest.mark.parametrize("which", ["ausm", "rusanov"])
def test_stretched_100_levels(solver, which):
    """1 deg grid shape: nAlts 100, stretched 1.5 -> 15 km, tidal winds."""
    spec = synth.GridSpec(2, 1, 100, stretched=True)
    cols = synth.make_columns(spec, seed=3)
    p = synth.StepParams()
    per_step(solver, cols, p, which, 3)


@pytest.mark.parametrize("n_alts", [10, 13, 17, 29, 60, 61, 124, 125, 188, 252])
def test_column_heights(solver, n_alts):
    """AUSM at column heights that change how the device forms split the
    column: the hydrostatic product's segments (two of 6 levels at nAlts 10,
    eight of 32 at 252, shorter last segments in between), the top
    boundary's lanes (in the first wavefront up to nAlts 60, across two at
    61, in the last one above), and one to four wavefronts per column; the
    Rusanov step too at three of them, for the boundary lanes."""
    spec = synth.GridSpec(1, 1, n_alts, stretched=n_alts >= 100)
    cols = subset(synth.make_columns(spec, seed=n_alts), [0, 7])
    per_step(solver, cols, synth.StepParams(), "ausm", 2)
    if n_alts in (10, 61, 252):  # the boundary lanes are the Rusanov step's too
        per_step(solver, cols, synth.StepParams(), "rusanov", 2)


@pytest.mark.parametrize("which", ["ausm", "rusanov"])
@pytest.mark.parametrize("improved", [True, False])
def test_ions_stretched_100_levels(solver, which, improved):
    """Ion advection (UseIonAdvection) on 100 stretched levels, 3 steps."""
    spec = synth.GridSpec(2, 1, 100, stretched=True)
    cols = synth.make_columns(spec, seed=15, ions=True)
    p = synth.StepParams(use_ion_advection=True, use_improved_ion_advection=improved)
    per_step(solver, cols, p, which, 3)


def test_full_1deg_grid_sampled(solver):
    """The full 64800 x 100 production grid, one step on the GPU; 96 sampled columns
    checked against the oracle (columns never interact), every column checked
    for finiteness and positivity, and a second run checked for bitwise
    reproducibility."""
    from gitm_amd import vertical
    spec = synth.CONFIGS["1deg_64800x100"]
    cols = synth.make_columns(spec, seed=11)
    p = synth.StepParams()
    d = vertical.to_device(cols)
    solver.advance(d, p, "ausm", 1)
    out = vertical.to_host(d)
    for k in ("nds", "temp", "rho", "nden"):
        assert np.all(out[k] > 0) and np.all(np.isfinite(out[k])), k
    for k in ("vvel", "vel", "logns"):
        assert np.all(np.isfinite(out[k])), k
    idx = np.random.default_rng(5).choice(spec.ncols, 96, replace=False)
    ref = cpu.advance(subset(cols, idx), p, solver=1, n_steps=1, threads=THREADS)
    assert_parity(subset(out, idx), ref, RTOL, what="1deg sampled")
    d2 = vertical.to_device(cols)
    solver.advance(d2, p, "ausm", 1)
    out2 = vertical.to_host(d2)
    for k in ("nds", "vvel", "vel", "temp", "rho"):
        assert np.array_equal(out[k], out2[k]), f"run-to-run difference in {k}"


@pytest.mark.parametrize("flags", [dict(use_neutral_friction=False),
                                   dict(use_coriolis=False),
                                   dict(dt=5.0, max_vert_vel=20.0),
                                   dict(beta_limiter=1.0)])
def test_flags(solver, flags):
    cols = subset(synth.make_columns(synth.CONFIGS["cpu_324x50"], seed=4),
                  np.arange(0, 324, 5))
    p = dataclasses.replace(synth.StepParams(), **flags)
    for which in ("ausm", "rusanov"):
        per_step(solver, cols, p, which, 2)


def test_tiles_are_bitwise_identical():
    """Every waves-per-column tile that holds the column gives the same bits
    (2 and 3 waves keep the RK4 carries in LDS, 4 in device memory)."""
    from gitm_amd import vertical
    cols = synth.make_columns(synth.GridSpec(1, 1, 100, stretched=True), seed=6)
    p = synth.StepParams()
    res = []
    for w in (2, 3, 4):
        s = vertical.VerticalSolver(waves_per_column=w)
        res.append(gpu_run(s, cols, p, "ausm", 2))
        s.close()
    for r in res[1:]:
        for k in ("nds", "vvel", "vel", "temp", "rho"):
            assert np.array_equal(r[k], res[0][k]), k


@pytest.mark.parametrize("ions", [False, True])
def test_host_entry_matches_device_entry(solver, ions):
    """gitm_vert_advance_host (host arrays staged by the library) gives the
    same bits as gitm_vert_advance on device tensors."""
    cols = subset(synth.make_columns(synth.CONFIGS["cpu_324x50"], seed=8, ions=ions),
                  np.arange(0, 324, 7))
    p = synth.StepParams(use_ion_advection=ions)
    dev = gpu_run(solver, cols, p, "ausm", 2)
    host = {k: np.array(v) for k, v in cols.items()}
    host["logns"] = np.zeros_like(host["nds"])
    host["nden"] = np.zeros_like(host["temp"])
    solver.advance_host(host, p, "ausm", 2)
    for k in ("nds", "vvel", "vel", "temp", "rho", "logns", "nden") + (("ions",) if ions else ()):
        assert np.array_equal(host[k], dev[k]), k


def test_nonpositive_density_is_reported(solver):
    """A column entering with NDensityS <= 0 stops the step (the reference
    calls stop_gitm, advance_vertical.f90:96-100); the error names it."""
    from gitm_amd import vertical
    cols = subset(synth.make_columns(synth.CONFIGS["cpu_324x50"], seed=9), np.arange(20))
    cols["nds"][13, 2, 30] = 0.0
    cols["nds"][17, 0, 5] = -1.0
    d = vertical.to_device(cols)
    with pytest.raises(vertical.VerticalStepError) as ei:
        solver.advance(d, synth.StepParams(), "ausm", 1)
    assert ei.value.code == 1 and ei.value.column == 13 and ei.value.level == 29
    out = vertical.to_host(d)
    assert np.array_equal(out["nds"][13], cols["nds"][13])  # left untouched
    solver.check()  # the error word was cleared


@pytest.mark.parametrize("which", ["ausm", "rusanov"])
def test_msis_lower_boundary(solver, which):
    """UseMSIS with the caller's level-0 MSIS values: per step against the
    oracle on 100 stretched levels, 5 steps, and the host-array entry point
    equal to the device one."""
    from gitm_amd import vertical
    cols = subset(synth.make_columns(synth.GridSpec(1, 1, 100, stretched=True), seed=41),
                  np.arange(0, 81, 3))
    cols["msis"] = synth.msis_bc_values(cols, seed=41)
    p = dataclasses.replace(synth.StepParams(), use_msis=True)
    sid = 1 if which == "ausm" else 0
    state = {k: np.array(v) for k, v in cols.items()}
    for step in range(5):
        out = gpu_run(solver, state, p, which)
        ref = cpu.advance(state, p, solver=sid, threads=THREADS)
        assert_parity(out, ref, RTOL, what=f"msis {which} step {step + 1}")
        state.update({k: ref[k] for k in ("nds", "vvel", "vel", "temp", "rho")})
    dev = gpu_run(solver, cols, p, which, n_steps=2)
    host = {k: np.array(v) for k, v in cols.items()}  # copies: advanced in place
    host["logns"] = np.zeros_like(host["nds"])
    host["nden"] = np.zeros_like(host["temp"])
    solver.advance_host(host, p, which, 2)
    for k in ("nds", "vvel", "vel", "temp", "rho"):
        assert np.array_equal(host[k], dev[k]), k


@pytest.mark.parametrize("which,flags", [
    ("ausm", {"use_gswm_tides": True}),
    ("rusanov", {"use_waccm_tides": True, "use_damping": True}),
    ("ausm", {"use_msis": True, "use_gswm_tides": True}),
    ("rusanov", {"use_damping": True})])
def test_tides_and_damping(solver, which, flags):
    """The bottom tides (set_vertical_bcs.Earth.f90:150-159; GSWM adds its
    temperature at every boundary pass, WACCM replaces it) from the caller's
    TidesEast/North/Temp, alone and after MSIS, and the Rusanov step's
    UseDamping VertTau profile (vertical_solver_rusanov.f90:581-584): per
    step against the oracle on 100 stretched levels, 5 steps, and the
    host-array entry point equal to the device one."""
    cols = subset(synth.make_columns(synth.GridSpec(1, 1, 100, stretched=True), seed=43),
                  np.arange(0, 81, 3))
    if flags.get("use_gswm_tides") or flags.get("use_waccm_tides"):
        cols["tides"] = synth.tides_values(cols, seed=43,
                                           absolute=bool(flags.get("use_waccm_tides")))
    if flags.get("use_msis"):
        cols["msis"] = synth.msis_bc_values(cols, seed=43)
    p = dataclasses.replace(synth.StepParams(), **flags)
    sid = 1 if which == "ausm" else 0
    state = {k: np.array(v) for k, v in cols.items()}
    for step in range(5):
        out = gpu_run(solver, state, p, which)
        ref = cpu.advance(state, p, solver=sid, threads=THREADS)
        assert_parity(out, ref, RTOL, what=f"{flags} {which} step {step + 1}")
        state.update({k: ref[k] for k in ("nds", "vvel", "vel", "temp", "rho")})
    dev = gpu_run(solver, cols, p, which, n_steps=2)
    host = {k: np.array(v) for k, v in cols.items()}
    host["logns"] = np.zeros_like(host["nds"])
    host["nden"] = np.zeros_like(host["temp"])
    solver.advance_host(host, p, which, 2)
    for k in ("nds", "vvel", "vel", "temp", "rho"):
        assert np.array_equal(host[k], dev[k]), k


def test_ausm_damping_and_bc_perturbation_are_unsupported(solver):
    """UseDamping with the AUSM solver and the user's BC perturbation return
    GITM_VERT_EUNSUPPORTED before any work, so the Fortran caller runs its
    own path; a tides flag without the tides' arrays is an argument error."""
    from gitm_amd import vertical
    cols = subset(synth.make_columns(synth.CONFIGS["cpu_324x50"], seed=44), np.arange(4))
    d = vertical.to_device(cols)
    for which, flags, code in (("ausm", {"use_damping": True}, 5),
                               ("rusanov", {"use_gswm_tides": True}, 4)):
        with pytest.raises(vertical.VerticalStepError) as ei:
            solver.advance(d, dataclasses.replace(synth.StepParams(), **flags), which, 1)
        assert ei.value.code == code, (which, flags, ei.value.code)
    assert np.array_equal(vertical.to_host(d)["temp"], cols["temp"])
