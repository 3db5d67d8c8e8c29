"""Writes the golden fixture of the horizontal step from the reference build.

The fixture holds, for one seeded synthetic block grid (two 9x9 blocks of
the 4x2-block 90-degree grid, 50 levels, Coriolis on, ion advection on with
GITM's default UseImprovedIonAdvection, so the ion-velocity divergence of
UAM_Gradient follows the step),
the interior cells the reference writes after advance_horizontal
(src/advance_horizontal.f90, compiled from the reference sources by ``make
-C oracle ref``) on six of the levels (each level is an independent 2-D
problem), and a SHA-256 of the block arrays it read, which the tests
regenerate from gitm_amd.synth.block_grid.  Run from the repository root:

    make -C oracle ref && python tests/golden/make_golden_horizontal.py
"""
from __future__ import annotations

import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

from gitm_amd import synth  # noqa: E402
from oracle import ref       # noqa: E402

HERE = os.path.dirname(os.path.abspath(__file__))
NAME = "horizontal_2blocks_ions.npz"
SPEC = synth.BlockGrid(4, 2, 50)
BLOCKS = (1, 6)
SEED = 7
OMEGA = 2.0 * np.pi / 86400.0
BETA = 2.0
LEVELS = (1, 2, 17, 33, 49, 50)


def digest(g: dict) -> str:
    import hashlib
    h = hashlib.sha256()
    for k in sorted(g):
        h.update(k.encode())
        h.update(np.ascontiguousarray(g[k], dtype=np.float64).tobytes())
    return h.hexdigest()


def inputs() -> tuple[dict, float]:
    full = synth.block_grid(SPEC, seed=SEED, ions=True)
    g = {k: np.ascontiguousarray(v[list(BLOCKS)]) for k, v in full.items()}
    return g, synth.horizontal_dt(g, cfl=0.4)


def main() -> None:
    g, dt = inputs()
    out = ref.horizontal(g, dt, OMEGA, BETA, use_coriolis=True, use_ion_advection=True,
                         use_improved_ion_advection=True)
    lv = np.array(LEVELS)
    keep = {"inputs_sha256": np.array(digest(g)), "dt": np.array(dt), "levels": lv}
    for k, v in out.items():  # interior cells of the chosen levels
        keep["out_" + k] = v[..., lv - 1, :, :] if k == "temp_source" else \
            v[..., lv + 1, 2:-2, 2:-2]
    np.savez_compressed(os.path.join(HERE, NAME), **keep)
    print(NAME, os.path.getsize(os.path.join(HERE, NAME)), "bytes, dt", dt)


if __name__ == "__main__":
    main()
