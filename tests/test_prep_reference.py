"""The oracle's column prep and vertical time step (ov_prepare) against the
reference build: calc_pressure (calc_pressure.f90:17-86), calc_rates'
MeanMajorMass and TempUnit (calc_rates.Earth.f90:34-50),
calc_eddy_diffusion_coefficient (Earth.f90:376-411) and
calc_timestep_vertical (calc_timestep.f90:85-147), compiled from the
reference sources (oracle/ref_prep_harness.f90).  Bit for bit: NDensity,
Pressure, cp, Gamma, MeanMajorMass, TempUnit, KappaEddyDiffusion, cMax(up)
and Dt, on the committed fixtures (tests/golden/prep_*.npz, written by
tests/golden/make_golden_prep.py) and, where the reference build is present,
on fresh seeded columns of four whole blocks; the reference's own Mass and
Vibration equal the defaults of synth.PrepParams."""
from __future__ import annotations

import dataclasses
import glob
import json
import os

import numpy as np
import pytest

from gitm_amd import synth
from oracle import cpu, ref

HERE = os.path.dirname(os.path.abspath(__file__))
FIXTURES = sorted(glob.glob(os.path.join(HERE, "golden", "prep_*.npz")))


def assert_same(out: dict, dt: float, want: dict, want_dt: float, what: str):
    for k in ref.PREP_OUT:
        a, b = out[k], want[k]
        if k == "cmax":  # computed on levels 0..nAlts+1 only
            a, b = a[:, 1:-1], b[:, 1:-1]
        assert np.array_equal(a, b), (what, k, np.max(np.abs(a - b)))
    assert dt == want_dt, (what, dt, want_dt)


def params_of(meta: dict) -> synth.PrepParams:
    return synth.PrepParams(**{k: tuple(v) if isinstance(v, list) else v
                               for k, v in meta["params"].items()})


def test_fixtures_present():
    names = {os.path.basename(f) for f in FIXTURES}
    assert {"prep_fixed_dt.npz", "prep_cfl_ions.npz", "prep_eddy_ramp_dt_end.npz"} <= names


@pytest.mark.parametrize("path", FIXTURES, ids=[os.path.basename(f) for f in FIXTURES])
def test_oracle_equals_reference_fixture(path):
    z = np.load(path)
    meta = json.loads(str(z["meta"]))
    pp = params_of(meta)
    cols = {k[3:]: z[k] for k in z.files if k.startswith("in_")}
    out, dt = cpu.prepare(cols, pp)
    want = {k: z["out_" + k] for k in ref.PREP_OUT}
    assert_same(out, dt, want, float(z["out_dt"]), os.path.basename(path))
    if "eddy_ramp" in path:  # the ramp is exercised, not only its ends
        e = want["eddy"]
        assert ((e > 0) & (e < pp.eddy_coef)).sum() >= 9 * 10
        assert float(z["out_dt"]) == pp.dt_end
    if "cfl" in path:  # the CFL step binds, not FixedDt
        assert float(z["out_dt"]) < pp.fixed_dt


def test_checker_sees_a_dropped_plus_one():
    """The bar catches a plausible slip: MeanMajorMass over NDensity instead
    of NDensity + 1 (calc_rates.Earth.f90:37) moves the mean mass."""
    z = np.load(FIXTURES[0])
    n = z["in_nds"].sum(axis=1)
    m = synth.species_masses()
    wrong = sum(m[s] * z["in_nds"][:, s] / n for s in range(6))
    assert not np.array_equal(wrong, z["out_mmm"])


needs_ref = pytest.mark.skipif(not ref.prep_available(),
                               reason="reference build not present (make -C oracle ref)")


@needs_ref
@pytest.mark.parametrize("change", [{}, {"fixed_dt": 1.0e9, "use_ion_advection": True},
                                    {"fixed_dt": 1.0e9, "eddy_p0": 3.0e-4, "eddy_p1": 3.0e-6,
                                     "cfl": 0.4}],
                         ids=["defaults", "cfl_ions", "ramp_cfl04"])
def test_oracle_equals_reference_fresh(change):
    pp = dataclasses.replace(synth.PrepParams(), **change)
    cols = synth.make_columns(synth.CONFIGS["cpu_324x50"], seed=77, ions=pp.use_ion_advection)
    cols = {k: np.ascontiguousarray(v) for k, v in cols.items()
            if k in ("nds", "temp", "rho", "vvel", "dalt", "ivel")}
    out_r, dt_r, mass, vib = ref.prep(cols, pp)
    out, dt = cpu.prepare(cols, pp)
    assert_same(out, dt, out_r, dt_r, str(change))
    assert np.array_equal(mass, np.asarray(synth.PrepParams().mass))
    assert tuple(vib) == tuple(synth.PrepParams().vibration)
