"""Batched ensembles (gitm_vert_batch_*): members with their own nAlts and
column counts stepped together.

Each member must come out exactly as the single-grid entry point leaves it
(the same column code runs, so bit for bit), and, through that, within the
parity bar of the oracle; per-member time steps from the batch prep must
equal the oracle's; a step is one launch, and its tiles must not change a
bit; a hipGraph-captured loop of prep + step must reproduce the eager loop
on every replay; an error must name its member and column.
"""
from __future__ import annotations

import os

import numpy as np
import pytest

from gitm_amd import synth
from oracle import cpu
from tests.helpers import assert_parity

pytestmark = pytest.mark.gpu

STATE = ("nds", "vvel", "vel", "temp", "rho", "ions")
THREADS = min(16, os.cpu_count() or 1)
# (nAlts, blocks of 9x9, exospheric temperature): four distinct nAlts, two
# of them shared, uneven column counts
LAYOUT = ((50, 1, 800.0), (120, 2, 1300.0), (60, 1, 700.0), (50, 2, 1100.0),
          (70, 1, 950.0), (120, 1, 1400.0))


def members(ions=False):
    ms = [synth.Member(synth.GridSpec(nb, 1, na), tinf, 500 + i)
          for i, (na, nb, tinf) in enumerate(LAYOUT)]
    return [synth.member_columns(m, ions=ions) for m in ms]


def single(solver, cols, params, which, n_steps=1):
    from gitm_amd import vertical
    d = vertical.to_device(cols)
    solver.advance(d, params, which, n_steps)
    return vertical.to_host(d)


def assert_same(out, ref, what):
    for k in STATE:
        if k in ref:
            assert np.array_equal(out[k], ref[k]), f"{what}: {k} differs"


@pytest.mark.parametrize("which,ions", [("ausm", False), ("rusanov", False), ("ausm", True)])
def test_batch_matches_single_and_oracle(solver, which, ions):
    from gitm_amd import vertical
    cols = members(ions)
    p = synth.StepParams(use_ion_advection=ions)
    dev = [vertical.to_device(c) for c in cols]
    b = vertical.VerticalBatch(solver, dev, ions=ions)
    assert b.n_alts == [50, 60, 70, 120]
    b.advance(p, which, 2)
    sid = 1 if which == "ausm" else 0
    for i, c in enumerate(cols):
        out = vertical.to_host(dev[i])
        assert_same(out, single(solver, c, p, which, 2), f"member {i}")
    for i in (0, 1):  # the smallest and the tallest grid against the oracle, one step
        d = vertical.to_device(cols[i])
        bb = vertical.VerticalBatch(solver, [d], ions=ions)
        bb.advance(p, which, 1)
        ref = cpu.advance(cols[i], p, solver=sid, threads=THREADS)
        errs = assert_parity(vertical.to_host(d), ref, 1e-11, what=f"member {i}")
        print(which, "ions" if ions else "", i, errs)
        bb.close()
    b.close()


def test_batch_prepare_per_member_dt(solver):
    """Each member gets its own CFL step, equal to the oracle's, and a step
    with those device steps equals single-grid steps with them."""
    import dataclasses

    import torch
    from gitm_amd import vertical
    cols = members()
    pp = dataclasses.replace(synth.PrepParams(), fixed_dt=1e9)
    dev = [vertical.to_device(c) for c in cols]
    b = vertical.VerticalBatch(solver, dev)
    dt = b.prepare(pp)
    got = dt.cpu().numpy()
    want = np.array([cpu.prepare(c, pp)[1] for c in cols])
    assert np.array_equal(got, want), (got, want)
    assert len(set(got.tolist())) > 1  # the members' steps differ
    # capped: every member at FixedDt
    dt2 = b.prepare(dataclasses.replace(pp, fixed_dt=0.25), dt_out=torch.empty_like(dt))
    assert np.all(dt2.cpu().numpy() == 0.25)
    # step with the per-member steps
    b.prepare(pp, dt_out=dt)
    b.advance(synth.StepParams(), "ausm", 1, dt=dt)
    for i, c in enumerate(cols):
        p_i = synth.StepParams(dt=float(got[i]))
        d = vertical.to_device(c)
        solver.prepare(d, pp)
        solver.advance(d, p_i, "ausm", 1)
        assert_same(vertical.to_host(dev[i]), vertical.to_host(d), f"member {i}")
    b.close()


def run_tiles(solver, cols, p, which, tiles):
    """Steps a fresh batch of ``cols`` once with each tile the library
    accepts; every member must equal its single-grid step bit for bit.
    Returns the tiles that ran."""
    from gitm_amd import vertical
    ref = [single(solver, c, p, which) for c in cols]
    ran = []
    for tile in tiles:
        dev = [vertical.to_device(c) for c in cols]
        b = vertical.VerticalBatch(solver, dev, ions=p.use_ion_advection)
        try:
            b.set_tile(tile)
        except vertical.VerticalStepError:
            b.close()
            continue
        assert b.tile() == tuple(tile)
        print("tile", tile, flush=True)
        b.advance(p, which, 1)
        for i in range(len(cols)):
            assert_same(vertical.to_host(dev[i]), ref[i], f"tile {tile} member {i}")
        b.close()
        ran.append(tuple(tile))
    return ran


def test_batch_every_tile_is_bitwise_identical(solver):
    """One launch per step under every tile of the search: workgroups of 2,
    3 or 4 wavefronts, 1..4 short columns per workgroup, carries in LDS or in
    device memory, one workgroup per task or a persistent grid."""
    from gitm_amd import vertical
    cols = members()
    b = vertical.VerticalBatch(solver, [vertical.to_device(c) for c in cols])
    tiles = b.candidate_tiles()
    assert b.tile() == (2, 2, 2, 0)  # the automatic tile
    with pytest.raises(vertical.VerticalStepError):
        b.set_tile((1, 1, 1, 0))  # 124 levels need 2 wavefronts
    with pytest.raises(vertical.VerticalStepError):
        b.set_tile((3, 2, 0, 1))  # a persistent grid needs every tall column to fill it
    b.close()
    ran = run_tiles(solver, cols, synth.StepParams(), "ausm", tiles)
    # 2 waves: 2 short column counts x 3 carry placements x 2 grids; 3 and
    # 4 waves: one task per workgroup, 3 + 4 short column counts x 3 carry
    # placements, less four 60-level columns with their carries in LDS
    # (172 KiB, more than a CU's 160)
    assert len(ran) == 32, ran
    assert (4, 4, 1, 0) not in ran and (4, 4, 2, 0) in ran
    assert any(t[3] == 1 for t in ran)


@pytest.mark.parametrize("which,ions", [("rusanov", False), ("ausm", True)])
def test_batch_tiles_other_paths(solver, which, ions):
    cols = members(ions)
    p = synth.StepParams(use_ion_advection=ions)
    ran = run_tiles(solver, cols, p, which,
                    [(2, 2, 1, 0), (2, 2, 0, 1), (2, 1, 0, 1), (4, 3, 0, 0), (2, 2, 2, 1)])
    assert len(ran) == 5


def test_batch_mixed_wavefront_counts(solver):
    """Tall members needing 2 and 3 wavefronts in one launch: the 2-wave
    columns leave a 3-wave workgroup's last wavefront empty (it ends at
    once), so only a grid of one workgroup per task is accepted."""
    from gitm_amd import vertical
    ms = [synth.Member(synth.GridSpec(1, 1, na), tinf, 700 + i)
          for i, (na, tinf) in enumerate(((150, 900.0), (70, 1200.0), (50, 800.0),
                                          (188, 1000.0)))]
    cols = [synth.member_columns(m) for m in ms]
    ran = run_tiles(solver, cols, synth.StepParams(), "ausm",
                    [(3, 3, 1, 0), (3, 1, 0, 0), (3, 3, 0, 1), (4, 4, 0, 0), (3, 3, 2, 0)])
    assert ran == [(3, 3, 1, 0), (3, 1, 0, 0), (4, 4, 0, 0), (3, 3, 2, 0)]


def test_batch_autotune(solver):
    from gitm_amd import vertical
    cols = members()
    p = synth.StepParams()
    ref = [single(solver, c, p, "ausm") for c in cols]
    dev = [vertical.to_device(c) for c in cols]
    b = vertical.VerticalBatch(solver, dev)
    times = b.autotune(p, "ausm", probe_columns=300)
    assert len(times) == 32 and all(t > 0 for t in times.values())
    assert b.tile() == min(times, key=times.get)
    b.advance(p, "ausm", 1)  # the probe runs leave the members untouched until now
    for i in range(len(cols)):
        assert_same(vertical.to_host(dev[i]), ref[i], f"tuned member {i}")
    b.close()


def test_batch_graph_replays(solver):
    """Three steps of prep + step captured in one graph, replayed three
    times from the same initial state with eager work and synchronisations
    between the replays: every replay equals the eager loop bit for bit."""
    import torch
    from gitm_amd import vertical
    cols = members()
    p = synth.StepParams()
    pp = synth.PrepParams()
    n_steps = 3
    # eager loop (also loads every kernel the capture will use)
    eager = [vertical.to_device(c) for c in cols]
    be = vertical.VerticalBatch(solver, eager)
    dte = torch.empty(len(cols), dtype=torch.float64, device="cuda")
    for _ in range(n_steps):
        be.prepare(pp, dt_out=dte)
        be.advance(p, "ausm", 1, dt=dte)
    want = [vertical.to_host(d) for d in eager]
    be.close()

    init = [vertical.to_device(c) for c in cols]
    dev = [{k: v.clone() for k, v in d.items()} for d in init]
    b = vertical.VerticalBatch(solver, dev)
    dt = torch.empty(len(cols), dtype=torch.float64, device="cuda")
    torch.cuda.synchronize()
    g = torch.cuda.CUDAGraph()
    with torch.cuda.graph(g):
        for _ in range(n_steps):
            b.prepare(pp, dt_out=dt)
            b.advance(p, "ausm", 1, dt=dt, check=False)
    for rep in range(3):
        for d, d0 in zip(dev, init):
            for k in d:
                d[k].copy_(d0[k])
        torch.cuda.synchronize()
        g.replay()
        torch.cuda.synchronize()
        b.check()
        for i in range(len(cols)):
            assert_same(vertical.to_host(dev[i]), want[i], f"replay {rep} member {i}")
    del g
    b.close()


def test_batch_error_names_member(solver):
    from gitm_amd import vertical
    cols = members()
    cols[3]["nds"][100, 2, 7] = 0.0  # member 3, its column 100, level 6
    dev = [vertical.to_device(c) for c in cols]
    b = vertical.VerticalBatch(solver, dev)
    with pytest.raises(vertical.VerticalStepError) as ei:
        b.advance(synth.StepParams(), "ausm", 1)
    assert ei.value.code == 1 and ei.value.member == 3
    assert ei.value.column == 100 and ei.value.level == 6
    # the failing column is left as it was; the other members advanced
    out = vertical.to_host(dev[3])
    assert np.array_equal(out["nds"][100], cols[3]["nds"][100])
    assert not np.array_equal(vertical.to_host(dev[0])["temp"], cols[0]["temp"])
    b.close()
