"""The oracle against the reference Fortran.

tests/golden/*.npz hold inputs and outputs of the reference vertical step
(compiled from the reference sources, oracle/Makefile `ref`, via
tests/golden/make_golden.py).  The C restatement must reproduce them bit for
bit.  When the reference build is present it is also run on fresh seeded
columns.
"""
from __future__ import annotations

import glob
import json
import os

import numpy as np
import pytest

from gitm_amd import synth
from oracle import cpu, ref

GOLDEN = sorted(g for g in glob.glob(os.path.join(os.path.dirname(__file__), "golden", "*.npz"))
                if not os.path.basename(g).startswith(("horizontal_", "sources_", "prep_")))
# (the vertical step's; horizontal_*, sources_* and prep_* are the other steps')
OUT = ("nds", "vvel", "vel", "temp", "rho", "logns", "nden", "ions")


def load_fixture(path):
    z = np.load(path)
    meta = json.loads(str(z["meta"]))
    cols = {k[3:]: z[k] for k in z.files if k.startswith("in_")}
    outs = {k[4:]: z[k] for k in z.files if k.startswith("out_")}
    return meta, cols, outs


def test_fixtures_present():
    assert len(GOLDEN) >= 9


@pytest.mark.parametrize("path", GOLDEN, ids=[os.path.basename(p)[:-4] for p in GOLDEN])
def test_oracle_bitwise_equals_reference(path):
    meta, cols, outs = load_fixture(path)
    params = synth.StepParams(**meta["params"])
    got = cpu.advance(cols, params, solver=meta["solver"], n_steps=meta["steps"])
    for k in OUT:
        if k in outs:
            assert np.array_equal(got[k], outs[k]), (
                f"{k}: max abs diff {np.max(np.abs(got[k] - outs[k]))}")


@pytest.mark.skipif(not ref.available(), reason="reference build (oracle/_ref) absent")
@pytest.mark.parametrize("solver", [0, 1])
@pytest.mark.parametrize("seed", [21, 22])
def test_oracle_bitwise_equals_reference_fresh(seed, solver):
    cols = synth.make_columns(synth.CONFIGS["cpu_324x50"], seed=seed, ions=True)
    idx = np.arange(seed % 5, 324, 23)
    cols = {k: np.ascontiguousarray(v[idx]) for k, v in cols.items()}
    params = synth.StepParams(use_ion_advection=(seed % 2 == 0))
    a = cpu.advance(cols, params, solver=solver, n_steps=4)
    b = ref.advance(cols, params, solver=solver, n_steps=4)
    for k in OUT:
        if k in b:
            assert np.array_equal(a[k], b[k]), k
