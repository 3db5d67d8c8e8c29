"""Writes the golden fixtures of the vertical step from the reference build.

Each fixture holds the inputs of a few seeded synthetic columns and the
outputs of the reference Fortran vertical step (advance_vertical ->
advance_vertical_1d_ausm / _rusanov, compiled from the reference sources by
``make -C oracle ref``) after n steps.  Run from the repository root:

    make -C oracle ref && python tests/golden/make_golden.py [name ...]
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

CASES = [
    # name, solver, steps, seed, columns (stride through the 324-column grid),
    # parameter overrides, generator overrides
    ("ausm_1step", 1, 1, 1, 27, {}, {}),
    ("rusanov_1step", 0, 1, 1, 27, {}, {}),
    ("ausm_10step", 1, 10, 2, 41, {}, {}),
    ("rusanov_10step", 0, 10, 2, 41, {}, {}),
    ("ausm_ions_3step", 1, 3, 3, 29, {"use_ion_advection": True}, {"ions": True}),
    ("rusanov_ions_3step", 0, 3, 3, 29, {"use_ion_advection": True}, {"ions": True}),
    ("ausm_plain_flags", 1, 2, 4, 53,
     {"use_neutral_friction": False, "use_coriolis": False}, {}),
    ("ausm_hot_dt5", 1, 2, 5, 47, {"dt": 5.0}, {"tinf": 1400.0}),
    ("ausm_ions_log", 1, 2, 6, 59,
     {"use_ion_advection": True, "use_improved_ion_advection": False}, {"ions": True}),
    # nighttime plasmasphere columns with a TEC below MinTEC keep their top
    # ion ghosts (set_vertical_bcs.Earth.f90:492-497,590)
    ("ausm_ions_low_tec", 1, 2, 7, 19, {"use_ion_advection": True}, {"ions": True}),
    # negative ion densities near the top are repaired before the top slope
    # (check_for_negative_densities, :608-649)
    ("rusanov_ions_negative_top", 0, 2, 8, 31, {"use_ion_advection": True},
     {"ions": True}),
    # the bottom tides (set_vertical_bcs.Earth.f90:150-159): GSWM adds its
    # temperature to Temp(-1:0) at every boundary pass, WACCM replaces it
    ("ausm_gswm_tides_3step", 1, 3, 9, 37, {"use_gswm_tides": True}, {}),
    ("rusanov_gswm_tides_3step", 0, 3, 9, 37, {"use_gswm_tides": True}, {}),
    ("ausm_ions_waccm_tides_2step", 1, 2, 12, 61,
     {"use_ion_advection": True, "use_waccm_tides": True}, {"ions": True}),
    # UseDamping: the Rusanov step's VertTau profile (vertical_solver_rusanov
    # .f90:581-584), with WACCM tides
    ("rusanov_damping_waccm_3step", 0, 3, 10, 43,
     {"use_damping": True, "use_waccm_tides": True}, {}),
    # UseMSIS through the reference's own msis_bcs (MSIS-00 GTD7, no HWM
    # winds): in_msis holds what it returned for each column's level 0, the
    # values the library takes from its caller
    ("ausm_msis_3step", 1, 3, 13, 23, {"use_msis": True}, {}),
    ("rusanov_msis_gswm_2step", 0, 2, 14, 29, {"use_msis": True, "use_gswm_tides": True}, {}),
]
# time and solar inputs of the UseMSIS cases (ModTime, ModInputs)
MSIS_ENV = dict(julian_day=172, utime=30000.0, f107=180.0, f107a=160.0)


def tweak(name: str, sub: dict, seed: int, pover: dict) -> None:
    """Edits of the generated columns that steer them into rare branches, and
    the tides' values of the cases that use them."""
    if pover.get("use_gswm_tides") or pover.get("use_waccm_tides"):
        sub["tides"] = synth.tides_values(sub, seed=seed,
                                          absolute=bool(pover.get("use_waccm_tides")))
    if pover.get("use_msis"):
        # each column's latitude (the sign seeded) and longitude, radians
        nc = sub["temp"].shape[0]
        z = synth.normal(seed, 305, (nc, 2))
        sub["lat"] = np.arccos(np.clip(sub["coslat"], -1.0, 1.0)) * np.sign(z[:, 0])
        sub["lon"] = np.mod(3.0 * z[:, 1], 2.0 * np.pi)
    if name == "ausm_ions_low_tec":
        sub["ions"] *= 1e-5           # TEC ~1e-4 TECU, far below MinTEC = 2
        sub["ions"][:, :9] = np.maximum(sub["ions"][:, :9], 1.0)
        sub["ions"][:, 9] = sub["ions"][:, :9].sum(axis=1)
        sub["sza"][:] = 2.5           # night ...
        sub["mlat"][::2] = 45.0       # ... half of them inside the plasmasphere band
    elif name == "rusanov_ions_negative_top":
        na = sub["temp"].shape[1] - 4
        top = na + 1                  # offset of level nAlts
        sub["ions"][0, 0, top - 1] = -5.0e3           # one negative level: averaged
        sub["ions"][1, 1, top - 2:top] = -1.0e3       # two in a row: 0.95 x below
        sub["ions"][2, 3, top + 1] = -2.0             # negative ghost at nAlts+1
        sub["ions"][3, 2, top - 4] = -7.0             # deepest repairable level


def main(only=None) -> None:
    if not ref.available():
        raise SystemExit("reference build missing: run `make -C oracle ref` first")
    spec = synth.CONFIGS["cpu_324x50"]
    for name, solver, steps, seed, stride, pover, gover in CASES:
        if only and name not in only:
            continue
        cols = synth.make_columns(spec, seed=seed, **gover)
        idx = np.arange(0, spec.ncols, stride)
        sub = {k: np.ascontiguousarray(v[idx]) for k, v in cols.items()}
        tweak(name, sub, seed, pover)
        params = dataclasses.replace(synth.StepParams(), **pover)
        env = MSIS_ENV if pover.get("use_msis") else None
        out = ref.advance(sub, params, solver=solver, n_steps=steps, msis_env=env)
        if env:
            sub["msis"] = out.pop("msis")
        meta = dict(solver=solver, steps=steps, seed=seed, columns=idx.tolist(),
                    msis_env=env,
                    params=dataclasses.asdict(params), generator=gover,
                    source="reference Fortran (oracle/_ref) via oracle/ref.py")
        arrays = {f"in_{k}": v for k, v in sub.items()}
        arrays.update({f"out_{k}": v for k, v in out.items()})
        path = os.path.join(HERE, f"{name}.npz")
        np.savez_compressed(path, meta=np.array(json.dumps(meta)), **arrays)
        print(f"{name}: {len(idx)} columns, {os.path.getsize(path) / 1024:.0f} KiB")


if __name__ == "__main__":
    main(sys.argv[1:])
