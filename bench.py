"""Throughput of the fused FP64 vertical step on one MI355X.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--solver ausm|rusanov]

Workload: the 1 deg x 1 deg production grid (180 x 360 = 64800 columns),
nAlts = 100 stretched 100-700 km, seeded synthetic atmosphere
(gitm_amd.synth, "1deg_64800x100"), dt = 2 s.  One step is one call of the
vertical step (src/advance_vertical.f90:8-39, one RK4 AUSM step per column)
over every column, i.e. one kernel launch.  Inputs are resident in HBM before
the timed region starts.

Prints one JSON line.  ``value`` is column-updates/s over the K timed steps.
``roofline`` prices the kernel's algorithmic HBM bytes per launch
(``bytes_per_column(nAlts)`` x columns) against its mean launch duration,
timed with HIP events on the stream the kernel is queued on; ``traffic`` is
the measured HBM bytes per launch when a counter summary is present under
profiles/.  ``cpu_baseline`` times the CPU restatement of the same step
(oracle/, bit-identical to the reference Fortran) on a bounded sample of the
same grid on this host's cores.

Further figures ride on the same line: ``prep`` (the column prep),
``sources`` (the sources step's conduction and viscosity on the same grid),
``ensemble`` (64 members with their own nAlts and column counts as one
batch, a hipGraph of 20 prep+step iterations; ``--no-ensemble`` skips it)
and ``adjoint`` with
``adjoint_forward_ratio`` (a 10-step window on the 5 deg grid, backward pass
with recompute over forward pass; ``--no-adjoint`` skips it), and
``cpu_reference``: the reference Fortran build on the host cores on the 5 deg
grid (its compiled nAlts), with the GPU's rate on the same columns.
"""
from __future__ import annotations

import argparse
import json
import os
import re
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

WORKLOAD = "1deg_64800x100"
HBM_PEAK_GBS = 8000.0
FP64_PEAK_TFLOPS = 78.6
TRAFFIC_FILE = os.path.join(ROOT, "profiles", "vert_column_ausm_1deg_64800x100_traffic.txt")


# FP64 operations of one AUSM column-update as the product column code does
# them (device forms; faces once; every lane of the column, idle tile lanes
# excluded), counted by class by running that code over a counting number
# type on the host (tests/emu/count.h; tests/test_flop_count.py recomputes
# them): column 0 of the seeded 1-degree workload (nAlts 100) and of the
# 5-degree one (nAlts 50).  Columns differ by about 0.1% (branches).
OPS_AUSM = {100: dict(add=351994, mul=540220, div=29130, sqrt=3232, trans=31057, cmp=137398),
            50: dict(add=178720, mul=273185, div=14880, sqrt=1632, trans=16157, cmp=68890)}
# weights: +, -, *, min, max, compare 1; divide and sqrt 10; exp, log, pow,
# tanh 20
OP_WEIGHTS = dict(add=1, mul=1, cmp=1, div=10, sqrt=10, trans=20)


def ops_per_column(n_alts: int, weighted: bool = True) -> float:
    """FP64 operations of one AUSM column-update (OPS_AUSM), weighted or raw."""
    c = OPS_AUSM[n_alts]
    return float(sum(c[k] * (OP_WEIGHTS[k] if weighted else 1) for k in c))


def bytes_per_column(n_alts: int) -> int:
    """Algorithmic HBM bytes of one column-update: 42 FP64 values per level.

    Reads (25 per level): log densities 6, species vertical winds 6, winds 3,
    temperature, log rho, MeanMajorMass, Gamma, eddy coefficient, cMax,
    altitude, gravity, 1/r, dAlt.  Writes (17): log densities 6, vertical
    winds 6, winds 3, temperature, log rho.
    """
    return 336 * (n_alts + 4)


def interface_bytes_per_column(n_alts: int) -> int:
    """HBM bytes the library's interface moves per column-update (GITM's own
    arrays, each value read or written once).

    Read and written on all nAlts+4 levels: 6 densities, 6 vertical winds,
    3 winds, temperature, rho (17 values).  Written only: 6 log densities and
    the total density (7).  Read only on all levels: TempUnit, MeanMajorMass,
    Gamma, altitude, gravity, 1/r, dAlt (7).  cMax on nAlts+2 levels, the eddy
    coefficient on nAlts levels, cos(latitude) once.
    """
    L = n_alts + 4
    return 8 * ((17 * 2 + 7 + 7) * L + (n_alts + 2) + n_alts + 1)


def prep_bytes_per_column(n_alts: int) -> int:
    """HBM bytes of the column prep (gitm_vert_prepare) per column: reads
    NDensityS (6), Temperature, Rho on all levels, the 6 vertical winds on
    0..nAlts+1 and dAlt on 1..nAlts; writes NDensity, Pressure, cp, Gamma,
    MeanMajorMass, TempUnit, the eddy coefficient on all levels and cMax on
    0..nAlts+1."""
    L = n_alts + 4
    return 8 * (8 * L + 6 * (n_alts + 2) + n_alts + 7 * L + (n_alts + 2))


def time_prep(torch, solver, dev, steps: int) -> dict:
    """Mean launch time of the column prep on the same grid (HIP events on
    the launch stream)."""
    from gitm_amd import synth
    for k in ("press", "cp"):
        dev[k] = torch.zeros_like(dev["temp"])
    dt = torch.empty(1, dtype=torch.float64, device="cuda")
    pp = synth.PrepParams()
    stream = torch.cuda.current_stream()
    solver.prepare(dev, pp, dt_out=dt)
    torch.cuda.synchronize()
    ev = [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True))
          for _ in range(steps)]
    for a, b in ev:
        a.record(stream)
        solver.prepare(dev, pp, dt_out=dt)
        b.record(stream)
    torch.cuda.synchronize()
    return {"launch_ms": float(np.mean([a.elapsed_time(b) for a, b in ev])),
            "dt": float(dt.item())}


def read_traffic(path: str):
    """HBM bytes per launch from the counter summary (FETCH_SIZE + WRITE_SIZE,
    scaled by the copy-kernel calibration), or None."""
    if not os.path.exists(path):
        return None
    m = re.search(r"^traffic_bytes_per_launch\s*=\s*([0-9.eE+]+)", open(path).read(), re.M)
    return float(m.group(1)) if m else None


def cpu_baseline(spec, solver: str, seconds: float, threads: int) -> dict:
    """Time the oracle (CPU restatement) on chunks of the same grid until
    ``seconds`` of wall time have passed."""
    from gitm_amd import synth
    from oracle import cpu

    cols = synth.make_columns(spec, seed=1)
    sid = cpu.AUSM if solver == "ausm" else cpu.RUSANOV
    chunk = 1024 * threads
    done = 0
    t0 = time.perf_counter()
    while done < spec.ncols:
        hi = min(done + chunk, spec.ncols)
        sub = {k: np.ascontiguousarray(v[done:hi]) for k, v in cols.items()}
        cpu.advance(sub, synth.StepParams(), sid, 1, threads=threads)
        done = hi
        if time.perf_counter() - t0 >= seconds:
            break
    el = time.perf_counter() - t0
    return {"value": done / el, "unit": "column-updates/s", "cores": threads,
            "kind": "port",
            "sample": f"first {done} of {spec.ncols} columns of the {spec.nalts}-level "
                      f"grid, one {solver.upper()} RK4 step each, {el:.1f} s wall on "
                      f"{threads} threads (C restatement, gcc -O2)"}


def cpu_reference(torch, solver, params, solver_name: str, threads: int,
                  n_steps: int = 200) -> dict | None:
    """The reference Fortran itself (oracle/_ref/gitm_vert_ref, compiled from
    its own sources, whose grid is nAlts = 50 as shipped) on the 5 deg grid
    (2592 columns x 50 levels): ``threads`` processes at once, each
    advancing its share of the columns ``n_steps`` steps, as the reference's
    MPI run spreads blocks over cores.  The wall time includes each process's
    file staging.  The GPU step is timed on the same columns beside it.
    None where the reference build is absent."""
    from concurrent.futures import ThreadPoolExecutor

    from gitm_amd import synth, vertical
    from oracle import cpu, ref

    if not ref.available():
        return None
    spec = synth.CONFIGS["5deg_2592x50"]
    cols = synth.make_columns(spec, seed=1)
    sid = cpu.AUSM if solver_name == "ausm" else cpu.RUSANOV
    parts = [{k: np.ascontiguousarray(v[i]) for k, v in cols.items()}
             for i in np.array_split(np.arange(spec.ncols), threads)]
    t0 = time.perf_counter()
    with ThreadPoolExecutor(threads) as ex:
        list(ex.map(lambda c: ref.advance(c, params, sid, n_steps), parts))
    el = time.perf_counter() - t0
    dev = vertical.to_device(cols, "cuda")
    for _ in range(3):
        solver.advance(dev, params, solver_name, 1, check=False)
    stream = torch.cuda.current_stream()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record(stream)
    for _ in range(n_steps):
        solver.advance(dev, params, solver_name, 1, check=False)
    e1.record(stream)
    torch.cuda.synchronize()
    solver.check()
    gpu = spec.ncols * n_steps / (e0.elapsed_time(e1) * 1e-3)
    return {"value": spec.ncols * n_steps / el, "unit": "column-updates/s", "cores": threads,
            "kind": "reference",
            "sample": f"all {spec.ncols} columns of the 5 deg grid (50 levels), {n_steps} "
                      f"{solver_name.upper()} RK4 steps, {threads} processes of the reference "
                      f"Fortran build, {el:.1f} s wall",
            "gpu_value_same_workload": gpu}


def copy_probe(torch, mib: int = 1024) -> None:
    """A plain device copy of known size, for calibrating FETCH_SIZE/WRITE_SIZE."""
    n = mib * (1 << 20) // 8
    a = torch.ones(n, dtype=torch.float64, device="cuda")
    b = torch.empty_like(a)
    b.copy_(a)
    torch.cuda.synchronize()
    print(f"copy probe: {mib} MiB read, {mib} MiB written", file=sys.stderr, flush=True)


def bench_ensemble(torch, solver, params, steps: int = 20, seed: int = 4) -> dict:
    """The ragged ensemble: 64 members with nAlts drawn from 50..120 and 2592,
    10368 or 25920 columns each (gitm_amd.synth.ensemble), advanced as one
    batch.  Each member's columns tile the first min(n, 2592) columns of its
    own seeded atmosphere.  The step's tile (workgroup size, short columns
    per workgroup, where the RK4 carries live, persistent grid or not) is
    autotuned on a probe batch; then a hipGraph holding ``steps`` x (batch
    prep with per-member CFL steps + batch step) is replayed once and timed.
    A batch step is one launch; the prep is two (the per-member step words
    are reset, then the columns' CFL limits folded in).  Also timed: the same
    steps at a fixed dt, one launch of ``steps`` steps."""
    from gitm_amd import synth, vertical

    mem = synth.ensemble(64, seed=seed)
    dev = []
    for m in mem:
        n = m.spec.ncols
        t = vertical.to_device(synth.member_columns(m, min(n, 2592)))
        reps = -(-n // t["temp"].shape[0])
        dev.append({k: v.repeat((reps,) + (1,) * (v.dim() - 1))[:n].contiguous()
                    for k, v in t.items()})
        del t
    ncols = sum(int(d["temp"].shape[0]) for d in dev)
    levels = sum(int(d["temp"].shape[0]) * (int(d["temp"].shape[1]) - 4) for d in dev)
    b = vertical.VerticalBatch(solver, dev)
    tune = b.autotune(params, "ausm")
    pp = synth.PrepParams()
    dt = torch.empty(len(dev), dtype=torch.float64, device="cuda")
    b.prepare(pp, dt_out=dt)
    b.advance(params, "ausm", 1, dt=dt)  # eager: loads every kernel before capture
    torch.cuda.synchronize()
    g = torch.cuda.CUDAGraph()
    with torch.cuda.graph(g):
        for _ in range(steps):
            b.prepare(pp, dt_out=dt)
            b.advance(params, "ausm", 1, dt=dt, check=False)
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    torch.cuda.synchronize()
    e0.record()
    g.replay()
    e1.record()
    torch.cuda.synchronize()
    b.check()
    graph_ms = e0.elapsed_time(e1) / steps
    e0.record()
    b.advance(params, "ausm", steps, check=False)
    e1.record()
    torch.cuda.synchronize()
    b.check()
    fixed_ms = e0.elapsed_time(e1) / steps
    ok = all(bool(torch.isfinite(d["temp"]).all()) and bool((d["nds"] > 0).all())
             for d in dev)
    tile = dict(zip(vertical.VerticalBatch.TILE_KEYS, b.tile()))
    del g
    b.close()
    return {"members": len(dev), "columns": ncols, "mean_n_alts": levels / ncols,
            "distinct_n_alts": sorted({int(d["temp"].shape[1]) - 4 for d in dev}),
            "tile": tile,
            "tiles_tried": {",".join(map(str, k)): round(v, 4) for k, v in tune.items()},
            "launches_per_step": {"step": 1, "prep": 2},
            "graph_steps": steps,
            "ms_per_step_graph_prep_and_step": graph_ms,
            "column_updates_per_s_graph": ncols / (graph_ms * 1e-3),
            "ms_per_step_fixed_dt": fixed_ms,
            "column_updates_per_s_fixed_dt": ncols / (fixed_ms * 1e-3),
            "state_finite_and_positive": ok}


def sources_bytes_per_column(n_alts: int) -> int:
    """Algorithmic HBM bytes of one column of the sources step (outputs on,
    state not updated): per level 31 FP64 values read (Temperature,
    TempUnit, Rho, eddy, cp, KappaTemp, ViscCoef, ViscCoefS 6, NDensityS 6,
    Velocity 3, VerticalVelocity 6, Altitude, RadialDistance, heating) and
    10 written (Conduction, Viscosity 3, VerticalViscosityS 6)."""
    return 41 * 8 * (n_alts + 4)


BLOCK_LON = BLOCK_LAT = 9   # GITM's blocks of 9x9 cells (ModSize.f90.orig)
NSPECIES_TOTAL = 11         # nSpeciesTotal: NDensityS holds them all; the step uses six


def block_layout(torch, dev: dict, n_alts: int):
    """The column-contiguous device columns ``dev`` in GITM's own block
    layout (ModGITM.f90:192-255): arrays (-1:nLons+2,-1:nLats+2,-1:nAlts+2
    [,nComp],nBlocks) of 9x9 blocks with their halo cells (NaN), NDensityS with
    all 11 species, cMax_GDB(0:nLons+1,0:nLats+1,0:nAlts+1,3,nBlocks),
    KappaEddyDiffusion without halos, CosLatitude(-1:nLats+2,nBlocks).
    Returns the device tensors (C order [block][comp][level][lat][lon]), the
    C-ABI descriptor and the (i, j, b) index of every column."""
    from gitm_amd import _abi
    ncols = int(dev["temp"].shape[0])
    nb, L = ncols // (BLOCK_LON * BLOCK_LAT), n_alts + 4
    NLo, NLg = BLOCK_LON + 4, BLOCK_LAT + 4
    c = torch.arange(ncols, device="cuda")
    i, j, b = c % BLOCK_LON, (c // BLOCK_LON) % BLOCK_LAT, c // (BLOCK_LON * BLOCK_LAT)
    nan = float("nan")

    def halo(key, nc_alloc=None):
        a = dev[key]
        a3 = a[:, None, :] if a.dim() == 2 else a
        out = torch.full((nb, nc_alloc or a3.shape[1], L, NLg, NLo), nan,
                         dtype=torch.float64, device="cuda")
        out[b, :a3.shape[1], :, j + 2, i + 2] = a3
        return out

    blk = {k: halo(k) for k in ("vvel", "vel", "temp", "rho", "gamma", "alt", "grav", "invr",
                                "dalt", "tempunit", "mmm", "logns", "nden")}
    blk["nds"] = halo("nds", NSPECIES_TOTAL)
    cm = torch.full((nb, 3, L - 2, BLOCK_LAT + 2, BLOCK_LON + 2), nan, dtype=torch.float64,
                    device="cuda")
    cm[b, 2, :, j + 1, i + 1] = dev["cmax"][:, 1:L - 1]
    blk["cmax"] = cm
    ed = torch.full((nb, L, BLOCK_LAT, BLOCK_LON), nan, dtype=torch.float64, device="cuda")
    ed[b, :, j, i] = dev["eddy"]
    blk["eddy"] = ed
    cl = torch.full((nb, NLg), nan, dtype=torch.float64, device="cuda")
    cl[b, j + 2] = dev["coslat"]
    blk["coslat"] = cl
    return blk, (i, j, b)


def block_columns(blk: dict, ptr_of, n_blocks: int, n_alts: int):
    """C-ABI descriptor of block-layout arrays at the addresses ptr_of(key)."""
    from gitm_amd import _abi
    c = _abi.Columns()
    c.n_lons, c.n_lats, c.n_blocks, c.n_alts = BLOCK_LON, BLOCK_LAT, n_blocks, n_alts

    def f(key, nc, **kw):
        return _abi.gitm_block_field(ptr_of(key), BLOCK_LON, BLOCK_LAT, n_alts, nc, **kw)

    c.ndensity_s = f("nds", NSPECIES_TOTAL)
    c.vertical_velocity, c.velocity, c.log_ns = f("vvel", 6), f("vel", 3), f("logns", 6)
    for key, name in (("temp", "temperature"), ("rho", "rho"), ("nden", "ndensity"),
                      ("tempunit", "temp_unit"), ("mmm", "mean_major_mass"),
                      ("gamma", "gamma"), ("alt", "altitude"), ("grav", "gravity"),
                      ("invr", "inv_radial_distance"), ("dalt", "dalt")):
        setattr(c, name, f(key, 1))
    c.cmax = f("cmax", 3, comp0=2, lon_lo=0, lat_lo=0, alt_lo=0, alt_hi=n_alts + 1)
    c.eddy = f("eddy", 1, lon_lo=1, lat_lo=1)
    c.cos_latitude = _abi.latitude_field(ptr_of("coslat"), BLOCK_LAT)
    return c


def bench_block_layout(torch, solver, dev, spec, params, which: str, reps: int = 10) -> dict:
    """The step on GITM's own block layout at 1 deg (800 blocks of 9x9, the
    layout the Fortran drop-in passes), device-resident: its time per launch
    beside the column-contiguous layout's, after a check that one step gives
    the same bits on both; and the host-array entry point on the same arrays
    in host memory (what gitm_vert_advance_host stages over PCIe per call)."""
    import ctypes

    from gitm_amd import vertical
    nb = spec.ncols // (BLOCK_LON * BLOCK_LAT)
    ref = {k: v.clone() for k, v in dev.items()}
    blk, (i, j, b) = block_layout(torch, ref, spec.nalts)
    c = block_columns(blk, lambda k: blk[k].data_ptr(), nb, spec.nalts)
    p = vertical.make_params(params)
    sid = vertical.solver_id(which)
    stream = torch.cuda.current_stream()
    solver.set_stream(stream.cuda_stream)

    def step():
        rc = vertical.lib().gitm_vert_advance(solver._h, ctypes.byref(p), ctypes.byref(c), sid, 1)
        if rc:
            raise RuntimeError(f"gitm_vert_advance on the block layout: {rc}")

    step()
    solver.advance(ref, params, which, 1, check=False)
    torch.cuda.synchronize()
    solver.check()
    same = all(torch.equal(blk[k][b, :ref[k].shape[1] if ref[k].dim() == 3 else 1, :,
                                  j + 2, i + 2].reshape(ref[k].shape), ref[k])
               for k in ("nds", "vvel", "vel", "temp", "rho"))
    ev = [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True))
          for _ in range(reps)]
    for e0, e1 in ev:
        e0.record(stream)
        step()
        e1.record(stream)
    torch.cuda.synchronize()
    solver.check()
    block_ms = float(np.mean([e0.elapsed_time(e1) for e0, e1 in ev]))
    for e0, e1 in ev:
        e0.record(stream)
        solver.advance(ref, params, which, 1, check=False)
        e1.record(stream)
    torch.cuda.synchronize()
    column_ms = float(np.mean([e0.elapsed_time(e1) for e0, e1 in ev]))
    # the host-array entry point on the block arrays in host memory
    host = {k: v.cpu().numpy() for k, v in blk.items()}
    del blk
    torch.cuda.empty_cache()
    ch = block_columns(host, lambda k: host[k].ctypes.data, nb, spec.nalts)
    staged = sum(v.nbytes for k, v in host.items() if k not in ("nds",)) + \
        host["nds"][:, :6].nbytes
    t = []
    for _ in range(3):
        t0 = time.perf_counter()
        rc = vertical.lib().gitm_vert_advance_host(solver._h, ctypes.byref(p), ctypes.byref(ch),
                                                   sid, 1)
        t.append(time.perf_counter() - t0)
        if rc:
            raise RuntimeError(f"gitm_vert_advance_host on the block layout: {rc}")
    return {"layout": "GITM blocks of 9x9 with halo cells, 800 blocks x 100 levels",
            "bitwise_equal_to_column_layout": bool(same),
            "launch_ms": block_ms, "column_layout_launch_ms": column_ms,
            "ratio_to_column_layout": block_ms / column_ms,
            "host_staged_ms_per_call": 1e3 * float(np.median(t[1:])),
            "host_staged_bytes_in_approx": staged}


def bench_sources(torch, solver, dev, spec, reps: int = 10) -> dict:
    """The sources step (gitm_vert_sources) on the same grid, increments
    only (the state is left as it is), timed with HIP events."""
    from gitm_amd import synth

    host = {k: dev[k].cpu().numpy() for k in ("nds", "temp", "tempunit", "alt", "invr")}
    src = {k: torch.from_numpy(v).cuda() for k, v in synth.source_fields(host).items()}
    src["conduction"] = torch.zeros_like(dev["temp"])
    src["viscosity"] = torch.zeros_like(dev["vel"])
    src["vert_viscosity"] = torch.zeros_like(dev["vvel"])
    solver.sources(dev, src, 2.0, apply=False)
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    torch.cuda.synchronize()
    e0.record()
    for _ in range(reps):
        solver.sources(dev, src, 2.0, apply=False)
    e1.record()
    torch.cuda.synchronize()
    ms = e0.elapsed_time(e1) / reps
    nbytes = sources_bytes_per_column(spec.nalts) * spec.ncols
    return {"kernel": "sources_kernel (gitm_vert_sources: calc_thermal_conduction, "
                      "calc_viscosity, add_sources)",
            "launch_ms": ms, "bound": "hbm", "bytes_per_launch": nbytes,
            "achieved": nbytes / (ms * 1e-3) / 1e9, "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": nbytes / (ms * 1e-3) / 1e9 / HBM_PEAK_GBS,
            "column_updates_per_s": spec.ncols / (ms * 1e-3)}


def horizontal_bytes_per_slab(nlo: int, nla: int, ions: bool = False) -> int:
    """Algorithmic HBM bytes of one altitude of one block of the horizontal
    step: every advected field read on the block's cells and halo cells
    ((n+4)^2 each: Rho, Temperature, Velocity 3, NDensityS 6,
    VerticalVelocity 6, with ions IDensityS 2) and written on its cells,
    plus HorizontalTempSource; Gamma, InvRadialDistance and
    RadialDistance on the cells; the face inverse distances on faces
    0..n+2 and the face distances on 1..n+1 of each row or column, the cell
    inverse distances, and cMax(east, north) on cells 0..n+1 of each row or
    column; the three latitude arrays."""
    nv = 17 + (2 if ions else 0)
    cells, halo = nlo * nla, (nlo + 4) * (nla + 4)
    reads = (nv * halo + 3 * cells + (nlo + 3) * nla + nlo * (nla + 3) + (nlo + 1) * nla
             + nlo * (nla + 1) + 2 * cells + (nlo + 2) * nla + nlo * (nla + 2) + 3 * nla)
    writes = (nv + 1) * cells
    return 8 * (reads + writes)


def bench_horizontal(torch, solver, reps: int = 10) -> dict:
    """The horizontal step (gitm_horiz_advance) on the 1-degree block grid
    (800 blocks of 9 x 9 cells, 100 levels), timed with HIP events."""
    from gitm_amd import synth

    spec = synth.BLOCK_CONFIGS["1deg_blocks_800x100"]
    host = synth.block_grid(spec, seed=1, topography=True)
    dt = synth.horizontal_dt(host, cfl=0.4)
    g = {k: torch.from_numpy(v).cuda() for k, v in host.items()}
    del host
    g["temp_source"] = torch.zeros_like(g["rho"])
    solver.horizontal(g, dt)
    torch.cuda.synchronize()
    solver.check()
    stream = torch.cuda.current_stream()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record(stream)
    for _ in range(reps):
        solver.horizontal(g, dt)
    e1.record(stream)
    torch.cuda.synchronize()
    solver.check()
    ms = e0.elapsed_time(e1) / reps
    # UseTopography: the same step with the levels advanced in order, one
    # launch of 800 workgroups per level
    solver.horizontal(g, dt, use_topography=True)
    e0.record(stream)
    for _ in range(reps):
        solver.horizontal(g, dt, use_topography=True)
    e1.record(stream)
    torch.cuda.synchronize()
    solver.check()
    tms = e0.elapsed_time(e1) / reps
    # the message exchange on the device (gitm_horiz_exchange), same grid
    g["itemp"], g["etemp"] = g["temp"].clone(), g["temp"].clone()
    solver.exchange(g, spec.nblocks_lon)
    e0.record(stream)
    for _ in range(reps):
        solver.exchange(g, spec.nblocks_lon)
    e1.record(stream)
    torch.cuda.synchronize()
    solver.check()
    xms = e0.elapsed_time(e1) / reps
    ok = bool(torch.isfinite(g["temp"]).all()) and bool((g["nds"] > 0).all())
    slabs = spec.nblocks * spec.nalts
    halo = (spec.nlons + 4) * (spec.nlats + 4) - spec.nlons * spec.nlats
    # 19 carried fields, one read and one write per halo cell and level
    xbytes = 2 * 8 * 19 * halo * spec.nalts * spec.nblocks
    nbytes = horizontal_bytes_per_slab(spec.nlons, spec.nlats) * slabs
    del g
    return {"kernel": "horiz_kernel<17, false> (gitm_horiz_advance: advance_horizontal, RK4 of "
                      "the Rusanov/MC lat-lon solver)",
            "workload": "1deg_blocks_800x100: 360x180 cells in 800 blocks of 9x9, "
                        "100 levels, Coriolis on, ions off",
            "dt": dt, "launch_ms": ms, "bound": "hbm", "bytes_per_launch": nbytes,
            "achieved": nbytes / (ms * 1e-3) / 1e9, "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": nbytes / (ms * 1e-3) / 1e9 / HBM_PEAK_GBS,
            "cell_updates_per_s": spec.ncells * spec.nalts / (ms * 1e-3),
            "topography": {"kernel": "horiz_kernel<17, true>, one launch per level",
                           "ms_per_step": tms,
                           "cell_updates_per_s": spec.ncells * spec.nalts / (tms * 1e-3)},
            "exchange": {"kernel": "exchange_kernel + exchange_check_kernel "
                                   "(gitm_horiz_exchange, neutrals)",
                         "launch_ms": xms, "halo_bytes": xbytes,
                         "achieved_gbs": xbytes / (xms * 1e-3) / 1e9},
            "state_finite_and_positive": ok}


def bench_adjoint(torch, solver, steps: int = 10, seed: int = 5) -> dict:
    """Differentiation workload: the 5 deg x 5 deg grid (2592 columns x 50
    levels), a window of ``steps`` x (column prep, AUSM step, Gaussian
    heating source peaked at 150 km, 1e-5 W/kg), loss = mean mass density at
    400 km.  The forward pass (the plain HIP prep and step, one checkpoint per
    step) and the backward pass (per step: the prep recomputed from the
    checkpoint, the hand-written adjoint of the step with the prep inputs'
    adjoints, the prep's adjoint) are timed over five passes after one
    untimed pass; each is the median of its five (a pass is tens of ms of
    small launches, whose time moves by several per cent from pass to pass).
    The ratio is the backward median over the forward median."""
    from gitm_amd import adjoint, synth, vertical

    spec = synth.CONFIGS["5deg_2592x50"]
    cols = synth.make_columns(spec, seed=seed)
    d = vertical.to_device(cols)
    for k in ("press", "cp"):
        d[k] = torch.zeros_like(d["temp"])
    solver.prepare(d, synth.PrepParams())
    fixed = {k: d[k] for k in adjoint.FIXED}
    fixed["cp"] = d["cp"]
    heat = torch.from_numpy(adjoint.heating_profile(cols["alt"])).cuda()
    w = torch.from_numpy(adjoint.level_weights(cols["alt"], 400e3)).cuda()
    p = synth.StepParams()
    fwd, bwd = [], []
    loss = 0.0
    for rep in range(6):
        state = {k: d[k].clone().requires_grad_(True) for k in adjoint.STATE}
        q = heat.clone().requires_grad_(True)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        out = adjoint.window(state, q, solver, fixed, p, steps)
        J = (out["rho"] * w).sum() / w.shape[0]
        torch.cuda.synchronize()
        t1 = time.perf_counter()
        J.backward()
        torch.cuda.synchronize()
        t2 = time.perf_counter()
        if rep:
            fwd.append((t1 - t0) * 1e3)
            bwd.append((t2 - t1) * 1e3)
        loss = float(J.detach())
    fwd_ms, bwd_ms = float(np.median(fwd)), float(np.median(bwd))
    ok = bool(torch.isfinite(state["temp"].grad).all()) and bool(torch.isfinite(q.grad).all())
    return {"workload": f"5deg_2592x50: 2592 columns x 50 levels, {steps} x (column prep + "
                        "AUSM step + heating source), loss = mean mass density at 400 km",
            "forward_ms": fwd_ms, "backward_ms": bwd_ms, "ratio": bwd_ms / fwd_ms,
            "forward_ms_per_step": fwd_ms / steps, "backward_ms_per_step": bwd_ms / steps,
            "forward_ms_samples": fwd, "backward_ms_samples": bwd,
            "loss": loss, "gradients_finite": ok}


def roofline_fp64(spec, launch_ms: float) -> dict:
    """The FP64 roofline of the same launch: the weighted operation count of
    the column code (OPS_AUSM) over the launch duration, against MI355X's
    78.6 TFLOP/s FP64 vector peak.  This is the binding bound: the step's
    intensity (about 56 weighted operations per algorithmic byte) is far
    above the FP64 ridge (78.6e12 / 8e12 = 9.8 per byte)."""
    ops = ops_per_column(spec.nalts) * spec.ncols
    raw = ops_per_column(spec.nalts, weighted=False) * spec.ncols
    achieved = ops / (launch_ms * 1e-3) / 1e12
    return {"bound": "fp64", "achieved": achieved, "peak": FP64_PEAK_TFLOPS,
            "unit": "TFLOP/s", "frac": achieved / FP64_PEAK_TFLOPS,
            "ops_per_launch": ops, "raw_ops_per_launch": raw,
            "weights": "+-*,min,max,cmp 1; div,sqrt 10; exp,log,pow,tanh 20",
            "intensity_ops_per_byte": ops / (bytes_per_column(spec.nalts) * spec.ncols),
            "binding": True}


def main() -> int:
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=50)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--solver", default="ausm", choices=("ausm", "rusanov"))
    ap.add_argument("--cpu-seconds", type=float, default=12.0,
                    help="wall-time budget of the CPU baseline sample (0 = skip)")
    ap.add_argument("--cpu-threads", type=int, default=0)
    ap.add_argument("--no-ensemble", action="store_true",
                    help="skip the 64-member ragged-ensemble figure")
    ap.add_argument("--no-horizontal", action="store_true",
                    help="skip the horizontal-step figure")
    ap.add_argument("--no-adjoint", action="store_true",
                    help="skip the adjoint/forward ratio")
    ap.add_argument("--copy-probe", action="store_true",
                    help="also run a 1 GiB device copy (counter calibration)")
    args = ap.parse_args()

    if args.gpus != 1:
        print(f"bench.py: this build runs on one GPU; --gpus {args.gpus} is not "
              "supported (columns are independent: N GPUs would each run their own "
              "share of the grid)", file=sys.stderr)
        return 2
    if args.steps < 1 or args.warmup < 0:
        print("bench.py: need --steps >= 1 and --warmup >= 0", file=sys.stderr)
        return 2

    import torch

    from gitm_amd import synth, vertical

    if not torch.cuda.is_available():
        print("bench.py: no GPU visible", file=sys.stderr)
        return 1
    spec = synth.CONFIGS[WORKLOAD]
    params = synth.StepParams()
    print(f"bench: {WORKLOAD} {args.solver} steps={args.steps} warmup={args.warmup}",
          file=sys.stderr, flush=True)
    cols = synth.make_columns(spec, seed=1)
    dev = vertical.to_device(cols, "cuda")
    del cols
    solver = vertical.VerticalSolver()
    stream = torch.cuda.current_stream()

    if args.copy_probe:
        copy_probe(torch)

    for _ in range(args.warmup):
        solver.advance(dev, params, args.solver, 1, check=False)
    torch.cuda.synchronize()
    solver.check()

    # Per-launch events on the launch stream (torch's current stream, which
    # VerticalSolver.advance hands to the library).
    ev = [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True))
          for _ in range(args.steps)]
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for a, b in ev:
        a.record(stream)
        solver.advance(dev, params, args.solver, 1, check=False)
        b.record(stream)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    solver.check()
    if not bool(torch.isfinite(dev["temp"]).all()) or not bool((dev["nds"] > 0).all()):
        print("bench.py: non-finite or non-positive state after the timed steps",
              file=sys.stderr)
        return 1
    launch_ms = float(np.mean([a.elapsed_time(b) for a, b in ev]))
    prep = time_prep(torch, solver, dev, max(args.steps, 10))
    src = bench_sources(torch, solver, dev, spec)
    print("bench: block layout", file=sys.stderr, flush=True)
    blk = bench_block_layout(torch, solver, dev, spec, params, args.solver)

    ncols = spec.ncols
    value = ncols * args.steps / wall
    per_launch = bytes_per_column(spec.nalts) * ncols
    achieved = per_launch / (launch_ms * 1e-3) / 1e9
    traffic = read_traffic(TRAFFIC_FILE) if args.solver == "ausm" else None

    hor = None
    if not args.no_horizontal:
        dev = None  # the 1-degree columns are no longer needed
        torch.cuda.empty_cache()
        print("bench: horizontal step", file=sys.stderr, flush=True)
        hor = bench_horizontal(torch, solver)
        torch.cuda.empty_cache()

    ens = None
    if not args.no_ensemble:
        dev = None
        torch.cuda.empty_cache()
        print("bench: 64-member ensemble", file=sys.stderr, flush=True)
        ens = bench_ensemble(torch, solver, params)
        torch.cuda.empty_cache()

    adj = None
    if not args.no_adjoint and args.solver == "ausm":
        print("bench: adjoint window", file=sys.stderr, flush=True)
        adj = bench_adjoint(torch, solver)
        torch.cuda.empty_cache()

    cpu = cpu_ref = None
    if args.cpu_seconds > 0:
        threads = args.cpu_threads or min(16, os.cpu_count() or 1)
        print(f"bench: CPU baseline on {threads} threads", file=sys.stderr, flush=True)
        cpu = cpu_baseline(spec, args.solver, args.cpu_seconds, threads)
        print(f"bench: reference Fortran on {threads} processes", file=sys.stderr, flush=True)
        cpu_ref = cpu_reference(torch, solver, params, args.solver, threads)

    line = {
        "metric": "FP64 AUSM RK4 column-updates/s @1°×1°×100 alts; adjoint/forward "
                  "time ratio",
        "value": value,
        "unit": "column-updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": wall * 1e3 / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (seeded hydrostatic atmosphere, gitm_amd.synth)",
        "config": {"workload": f"{WORKLOAD}: 1°×1° grid, 64800 columns x 100 levels "
                               f"(stretched 100-700 km), {args.solver} RK4, dt=2 s",
                   "n_columns": ncols, "n_alts": spec.nalts, "solver": args.solver},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": achieved / HBM_PEAK_GBS,
                     "traffic": traffic,
                     "algorithmic_bytes_per_launch": per_launch,
                     "interface_bytes_per_launch":
                         interface_bytes_per_column(spec.nalts) * ncols,
                     "launch_ms": launch_ms,
                     "kernel": f"vert_column_kernel<{64 * vertical.auto_waves(spec.nalts)},"
                               f"{1 if args.solver == 'ausm' else 0}>"},
        "roofline_fp64": roofline_fp64(spec, launch_ms) if args.solver == "ausm" else None,
        "cpu_baseline": cpu,
        "cpu_reference": cpu_ref,
        "prep": {"kernel": "prep_kernel (gitm_vert_prepare: calc_pressure, calc_rates "
                           "mean mass, eddy coefficient, calc_timestep_vertical)",
                 "launch_ms": prep["launch_ms"], "bound": "hbm",
                 "bytes_per_launch": prep_bytes_per_column(spec.nalts) * ncols,
                 "achieved": prep_bytes_per_column(spec.nalts) * ncols
                 / (prep["launch_ms"] * 1e-3) / 1e9,
                 "peak": HBM_PEAK_GBS, "unit": "GB/s",
                 "frac": prep_bytes_per_column(spec.nalts) * ncols
                 / (prep["launch_ms"] * 1e-3) / 1e9 / HBM_PEAK_GBS,
                 "cfl_dt": prep["dt"]},
        "sources": src,
        "block_layout": blk,
        "horizontal": hor,
        "ensemble": ens,
        "adjoint_forward_ratio": adj["ratio"] if adj else None,
        "adjoint": adj,
    }
    print(json.dumps(line), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
