"""GPU message exchange (gitm_horiz_exchange) against the oracle's
(oh_exchange, pinned by geometry in test_exchange_oracle.py): copies and
sign changes only, so bit for bit, on several block grids with and without
ions; then an exchange followed by a horizontal step on the 1-degree grid
sample matches the oracle's exchange and step."""
from __future__ import annotations

import numpy as np
import pytest

from gitm_amd import synth
from oracle import cpu
from tests.test_exchange_oracle import sphere_grid

pytestmark = pytest.mark.gpu

KEYS = cpu.EXCHANGED + ("ions", "ivel")


def gpu_exchange(solver, grid, nbl, ions):
    import torch
    d = {k: torch.from_numpy(np.ascontiguousarray(v)).cuda() for k, v in grid.items()}
    solver.exchange(d, nbl, use_ion_advection=ions, ion_velocity="ivel")
    torch.cuda.synchronize()
    return d, {k: v.cpu().numpy() for k, v in d.items()}


@pytest.mark.parametrize("nbl,nbt,ions", [(4, 3, True), (1, 2, False), (6, 1, True),
                                          (8, 4, False)])
def test_exchange_matches_oracle(solver, nbl, nbt, ions):
    grid, _ = sphere_grid(nbl, nbt, ions=True)
    nb, L, NLg, NLo = grid["rho"].shape
    hm = np.ones((NLg, NLo), bool)
    hm[2:-2, 2:-2] = False
    for k in KEYS:
        grid[k][..., 2:L - 2, :, :][..., hm] = 1.0e30
    print(f"exchange {nbl}x{nbt} ions={ions}", flush=True)
    _, out = gpu_exchange(solver, grid, nbl, ions)
    solver.check()
    ref, rc, _ = cpu.exchange(grid, nbl, ions)
    assert rc == 0
    for k in ref:  # without ions: the ion velocity only changes sign at the poles
        assert np.array_equal(out[k], ref[k]), k
    if not ions:
        assert np.array_equal(out["ions"], grid["ions"])


def test_exchange_reports_low_temperature(solver):
    from gitm_amd.vertical import VerticalStepError
    grid, _ = sphere_grid(2, 2, ions=False)
    grid["temp"][3, 2, 4, 5] = 50.0
    gpu_exchange(solver, grid, 2, False)
    with pytest.raises(VerticalStepError) as e:
        solver.check()
    assert e.value.code == 9 and e.value.level == 1 and e.value.column == (3 * 5 + 2) * 6 + 3


def test_exchange_then_step_on_one_degree_blocks(solver):
    """The eight blocks of the 1-degree grid nearest the north pole in two
    block rows, exchanged on the device as one 8-block-wide sub-grid would
    be if it were the sphere (longitudes wrap over the eight blocks), then
    stepped: equal to the oracle's exchange and step."""
    spec = synth.BlockGrid(8, 2, 30)
    g = synth.block_grid(spec, seed=9, ions=True)
    g["itemp"], g["etemp"] = g["temp"] * 1.1, g["temp"] * 1.3
    dt = synth.horizontal_dt(g, cfl=0.4)
    d, _ = gpu_exchange(solver, g, 8, True)
    solver.horizontal(d, dt, use_ion_advection=True, use_improved_ion_advection=True)
    import torch
    torch.cuda.synchronize()
    solver.check()
    ref, rc, _ = cpu.exchange(g, 8, True)
    assert rc == 0
    h = dict(g)
    h.update(ref)
    ref2, rc, _ = cpu.horizontal(h, dt, use_ion_advection=True, use_improved_ion_advection=True)
    assert rc == 0
    for k in ("rho", "temp", "vel", "vvel", "nds", "ions"):
        assert np.array_equal(d[k].cpu().numpy(), ref2[k]), k
