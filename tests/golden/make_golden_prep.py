"""Writes the golden fixtures of the column prep and the vertical time step
from the reference build.

Each fixture holds nine seeded synthetic columns (nAlts 50; with ion
advection also their ion winds), run as one 9x9 block of them, the prep
parameters, and the outputs
of the reference's calc_pressure, calc_rates, calc_eddy_diffusion_coefficient
and calc_timestep_vertical (src/calc_pressure.f90, src/calc_rates.Earth.f90,
src/Earth.f90, src/calc_timestep.f90, compiled from the reference sources by
``make -C oracle ref``): NDensity, Pressure, cp, Gamma, MeanMajorMass,
TempUnit, KappaEddyDiffusion, cMax(up) and Dt.  Run from the repository
root:

    make -C oracle ref && python tests/golden/make_golden_prep.py
"""
from __future__ import annotations

import dataclasses
import json
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

from gitm_amd import synth  # noqa: E402
from oracle import ref       # noqa: E402

HERE = os.path.dirname(os.path.abspath(__file__))
INPUTS = ("nds", "temp", "rho", "vvel", "dalt", "ivel")
CASES = [
    # name, seed, prep parameters changed from the defaults
    ("prep_fixed_dt", 41, {}),
    ("prep_cfl_ions", 42, {"fixed_dt": 1.0e9, "use_ion_advection": True}),
    # the eddy ramp between its pressures across levels 3..16, DtEnd's cap
    ("prep_eddy_ramp_dt_end", 43, {"fixed_dt": 1.0e9, "dt_end": 0.5, "eddy_p0": 1.0e-3,
                                   "eddy_p1": 1.0e-5}),
]


def main() -> None:
    if not ref.prep_available():
        raise SystemExit("reference build missing: run `make -C oracle ref` first")
    spec = synth.CONFIGS["cpu_324x50"]
    for name, seed, change in CASES:
        pp = dataclasses.replace(synth.PrepParams(), **change)
        cols = synth.make_columns(spec, seed=seed, ions=pp.use_ion_advection)
        # nine distinct columns, each filling a row of the 9x9 block: the
        # block's outputs are those of the nine, and so is its time step
        sub = {k: np.ascontiguousarray(cols[k][:9]) for k in INPUTS if k in cols}
        if "ivel" not in sub:
            sub["ivel"] = np.zeros((9, 3, sub["temp"].shape[1]))
        block = {k: np.concatenate([v] * 9) for k, v in sub.items()}
        out, dt, _, _ = ref.prep(block, pp)
        for k in ref.PREP_OUT:
            assert np.array_equal(out[k], np.concatenate([out[k][:9]] * 9)), k
        meta = dict(seed=seed, columns=9, params=dataclasses.asdict(pp),
                    source="reference Fortran (oracle/_ref/gitm_prep_ref) via oracle/ref.py, "
                           "the nine columns repeated over one 9x9 block")
        arrays = {f"in_{k}": sub[k] for k in INPUTS}
        arrays.update({f"out_{k}": out[k][:9] for k in ref.PREP_OUT})
        arrays["out_dt"] = np.array(dt)
        path = os.path.join(HERE, f"{name}.npz")
        np.savez_compressed(path, meta=np.array(json.dumps(meta)), **arrays)
        print(f"{name}: dt {dt!r}, {os.path.getsize(path) / 1024:.0f} KiB")


if __name__ == "__main__":
    main()
