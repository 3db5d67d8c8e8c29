"""The frozen FP64 operation counts of bench.py (OPS_AUSM) are what the
product column code does: recounted here by running that code (host build,
device forms) over a counting number type (tests/emu/count.h) on the same
columns.  And a sanity check of the counter on the limiter."""
from __future__ import annotations

import numpy as np
import pytest

import bench
from gitm_amd import synth


@pytest.mark.parametrize("workload,n_alts", [("1deg_64800x100", 100), ("5deg_2592x50", 50)])
def test_frozen_counts(workload, n_alts):
    from tests.emu import emu
    cols = synth.make_columns(synth.CONFIGS[workload], seed=1)
    cols = {k: np.ascontiguousarray(v[:1]) for k, v in cols.items()}
    got = emu.count_ops(cols, synth.StepParams(), 1, 0, device_math=True)
    assert got == bench.OPS_AUSM[n_alts]


def test_count_is_near_the_estimate():
    # the estimate made by counting the reference's operations by hand:
    # 2.0e4 (+-35%) weighted operations per level per step
    per_level = bench.ops_per_column(100) / 104
    assert 1.3e4 < per_level < 2.7e4
    # the reference forms (more logs and divisions) count more
    from tests.emu import emu
    cols = synth.make_columns(synth.CONFIGS["5deg_2592x50"], seed=1)
    cols = {k: np.ascontiguousarray(v[:1]) for k, v in cols.items()}
    ref = emu.count_ops(cols, synth.StepParams(), 1, 0, device_math=False)
    dev = bench.OPS_AUSM[50]
    assert ref["trans"] > dev["trans"] and ref["div"] > dev["div"]
