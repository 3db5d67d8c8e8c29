"""Writes the golden fixtures of the sources step's thermal conduction and
viscosity from the reference build.

Each fixture holds one 9x9 block of seeded synthetic columns (nAlts 50) with
the sources step's coefficients (gitm_amd.synth.source_fields; ViscCoef and
ViscCoefS as the reference's calc_viscosity_coef forms them) and the outputs
of the reference's calc_thermal_conduction and calc_viscosity
(src/calc_conduction.f90, src/calc_viscosity.f90, compiled from the
reference sources by ``make -C oracle ref``): Conduction, Viscosity and
VerticalViscosityS on levels 0..nAlts+1.  Run from the repository root:

    make -C oracle ref && python tests/golden/make_golden_sources.py
"""
from __future__ import annotations

import json
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

from gitm_amd import synth  # noqa: E402
from oracle import ref       # noqa: E402

HERE = os.path.dirname(os.path.abspath(__file__))
INPUTS = ("temp", "tempunit", "rho", "cp", "eddy", "nds", "kappa_temp", "vel", "vvel",
          "alt", "raddist", "mmm", "visc_coef", "visc_coef_s")
CASES = [
    # name, seed, UseTurbulentCond, Dt
    ("sources_laminar", 31, False, 2.0),
    ("sources_turbulent_dt5", 32, True, 5.0),
]


def main() -> None:
    if not ref.sources_available():
        raise SystemExit("reference build missing: run `make -C oracle ref` first")
    spec = synth.CONFIGS["cpu_324x50"]
    for name, seed, turbulent, dt in CASES:
        cols = synth.make_columns(spec, seed=seed)
        cols.update(synth.source_fields(cols, seed=seed))
        sub = {k: np.ascontiguousarray(v[:81]) for k, v in cols.items()}  # one block
        out = ref.sources(sub, dt, turbulent)
        sub["visc_coef"], sub["visc_coef_s"] = out["visc_coef"], out["visc_coef_s"]
        meta = dict(seed=seed, use_turbulent_cond=turbulent, dt=dt, columns=81,
                    source="reference Fortran (oracle/_ref/gitm_sources_ref) via "
                           "oracle/ref.py")
        arrays = {f"in_{k}": sub[k] for k in INPUTS}
        arrays.update({f"out_{k}": out[k] for k in ("conduction", "viscosity",
                                                     "vert_viscosity")})
        path = os.path.join(HERE, f"{name}.npz")
        np.savez_compressed(path, meta=np.array(json.dumps(meta)), **arrays)
        print(f"{name}: {os.path.getsize(path) / 1024:.0f} KiB")


if __name__ == "__main__":
    main()
