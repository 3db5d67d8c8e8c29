"""Writes the golden fixtures of the vertical step from the reference build.

Each fixture holds the inputs of a few seeded synthetic columns and the
outputs of the reference Fortran vertical step (advance_vertical ->
advance_vertical_1d_ausm / _rusanov, compiled from the reference sources by
``make -C oracle ref``) after n steps.  Run from the repository root:

    make -C oracle ref && python tests/golden/make_golden.py
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
]


def main() -> None:
    if not ref.available():
        raise SystemExit("reference build missing: run `make -C oracle ref` first")
    spec = synth.CONFIGS["cpu_324x50"]
    for name, solver, steps, seed, stride, pover, gover in CASES:
        cols = synth.make_columns(spec, seed=seed, **gover)
        idx = np.arange(0, spec.ncols, stride)
        sub = {k: np.ascontiguousarray(v[idx]) for k, v in cols.items()}
        params = dataclasses.replace(synth.StepParams(), **pover)
        out = ref.advance(sub, params, solver=solver, n_steps=steps)
        meta = dict(solver=solver, steps=steps, seed=seed, columns=idx.tolist(),
                    params=dataclasses.asdict(params), generator=gover,
                    source="reference Fortran (oracle/_ref) via oracle/ref.py")
        arrays = {f"in_{k}": v for k, v in sub.items()}
        arrays.update({f"out_{k}": v for k, v in out.items()})
        path = os.path.join(HERE, f"{name}.npz")
        np.savez_compressed(path, meta=np.array(json.dumps(meta)), **arrays)
        print(f"{name}: {len(idx)} columns, {os.path.getsize(path) / 1024:.0f} KiB")


if __name__ == "__main__":
    main()
