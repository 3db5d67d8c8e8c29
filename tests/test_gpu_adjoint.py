"""GPU adjoint of the AUSM step and of the column prep, and the
differentiable window (gitm_amd.adjoint.VerticalWindow).

* The hand-written adjoint (gitm_vert_adjoint) against the taped adjoint
  (gitm_vert_adjoint_taped, every operation on a tape) on the same GPU, and
  against the CPU FP64 reference gradient: the host build of the taped
  column code (tests/emu), itself checked against central differences of
  the oracle in test_adjoint_cpu.py.  Bar: 1e-12 normwise per field.
* The adjoints of the column-prep inputs (gitm_vert_adjoint_all) against the
  host build of the same code (whose values test_adjoint_cpu.py pins to
  central differences of the oracle), and the adjoint of the column prep
  (gitm_vert_prepare_adjoint) against central differences of the GPU prep.
* The 10-step window (column prep, AUSM step, heating source; the loss the
  mean mass density at 400 km) on the full 5-degree grid (2592 columns x 50
  levels): on 32 random cells of the initial temperature and of the heating
  rate, the adjoint against central differences, each cell's difference
  taken on its own column's term of the loss (columns are independent, so
  one cell moves only its column's term).  Bar 1e-6 relative.  Temperature
  cells step by h = 1e-6 relative.  The heating (1e-5 W/kg at its 150 km
  peak) is ~0 at the levels the loss is sensitive to, so its cells step by
  an absolute 0.01 W/kg.  Cells are drawn from those whose term moves by
  more than 1e-8 of itself under the step, so that the differences' own
  rounding (about 1e-15 of the term) stays near 1e-7 of the derivative.
  The loss is piecewise smooth (limiters, upwinding, the maximum in cMax,
  the eddy coefficient's ramp): where the central difference misses (a
  switch within the step), the adjoint must match one one-sided difference
  (the smooth side) to 2e-6 instead; at most 4 of 32 cells may.
* The window with the prep held (with_prep=False) against the same window
  on the CPU: the oracle forward and, per step backwards, the host build of
  the taped adjoint.  Bar 1e-9 normwise: the GPU's forward trajectory
  differs from the oracle's by the last bits of exp/log, amplified over the
  10 steps.
"""
from __future__ import annotations

import ctypes

import numpy as np
import pytest

from gitm_amd import synth
from tests.helpers import subset

pytestmark = pytest.mark.gpu

STATE = ("nds", "vvel", "vel", "temp", "rho")
PREP = ("tempunit", "gamma", "mmm", "cmax", "eddy")


def _gout(cols, seed):
    rng = np.random.default_rng(seed)
    return {k: rng.standard_normal(cols[k].shape) / np.max(np.abs(cols[k])) for k in STATE}


def _normwise(a, b):
    return float(np.max(np.abs(a - b)) / np.max(np.abs(b)))


def test_hand_adjoint_matches_taped_on_gpu(solver):
    """324 columns of the 2x2-block grid: hand-written vs taped, one step."""
    import torch
    from gitm_amd import adjoint, vertical
    cols = synth.make_columns(synth.CONFIGS["cpu_324x50"], seed=3)
    p = synth.StepParams()
    d = vertical.to_device(cols)
    go = {k: torch.from_numpy(v).cuda() for k, v in _gout(cols, 11).items()}
    hand = adjoint.adjoint_step(solver, d, go, p)
    solver.check()
    taped = adjoint.adjoint_step(solver, d, go, p, taped=True)
    for k in STATE:
        err = _normwise(hand[k].cpu().numpy(), taped[k].cpu().numpy())
        print(k, "normwise", err)
        assert err < 1e-12, (k, err)
    # in place (gin = gout) gives the same bits
    go2 = {k: v.clone() for k, v in go.items()}
    gi2 = adjoint.adjoint_step(solver, d, go2, p, gin=go2)
    for k in STATE:
        assert torch.equal(gi2[k], hand[k]), k


def test_hand_adjoint_matches_cpu_taped_build(solver):
    import torch
    from gitm_amd import adjoint, vertical
    from tests.emu import emu
    cols = subset(synth.make_columns(synth.CONFIGS["cpu_324x50"], seed=3), [0, 17, 200])
    p = synth.StepParams()
    gout = _gout(cols, 12)
    ref, _, _ = emu.adjoint(cols, p, gout)
    d = vertical.to_device(cols)
    go = {k: torch.from_numpy(gout[k]).cuda() for k in STATE}
    gi = adjoint.adjoint_step(solver, d, go, p)
    for k in STATE:
        err = _normwise(gi[k].cpu().numpy(), ref[k])
        assert err < 1e-12, (k, err)


def test_prep_input_adjoints_match_host_build(solver):
    import torch
    from gitm_amd import adjoint, vertical
    from tests.emu import emu
    cols = subset(synth.make_columns(synth.CONFIGS["cpu_324x50"], seed=6), [1, 40, 99, 300])
    p = synth.StepParams()
    gout = _gout(cols, 13)
    gi_ref, gf_ref = emu.adjoint_hand(cols, p, gout, fixed=True, device_math=True)
    d = vertical.to_device(cols)
    go = {k: torch.from_numpy(gout[k]).cuda() for k in STATE}
    gprep = {k: torch.zeros_like(d["temp"]) for k in PREP}
    gi = adjoint.adjoint_step(solver, d, go, p, gprep=gprep)
    solver.check()
    for k in STATE:
        assert _normwise(gi[k].cpu().numpy(), gi_ref[k]) < 1e-12, k
    for k in PREP:
        ref = gf_ref[k]
        if k == "cmax":  # levels 0..nAlts+1
            ref, got = ref[:, 1:-1], gprep[k].cpu().numpy()[:, 1:-1]
        else:
            got = gprep[k].cpu().numpy()
        err = _normwise(got, ref)
        print(k, err)
        assert err < 1e-11, (k, err)


def test_prep_adjoint_against_differences(solver):
    """The adjoint of the column prep against central differences of the GPU
    prep, for a random linear functional of its outputs."""
    import torch
    from gitm_amd import adjoint, vertical
    cols = subset(synth.make_columns(synth.CONFIGS["cpu_324x50"], seed=7), np.arange(0, 324, 27))
    d = vertical.to_device(cols)
    geometry = {k: d[k] for k in adjoint.GEOMETRY}
    state = {k: d[k] for k in STATE}
    # a wide eddy-coefficient ramp (EddyDiffusionPressure0/1), so that the
    # coefficient depends on Pressure = Temperature Rho on many levels
    pp = synth.PrepParams(eddy_p0=1.0, eddy_p1=1.0e-4)
    rng = np.random.default_rng(4)
    base = adjoint._prep(solver, state, geometry, pp)
    # the functional's weights: random, scaled by each output's size
    wn = {}
    for k in PREP + ("cp",):
        w = torch.from_numpy(rng.standard_normal(d["temp"].shape)).cuda()
        o = base[k][:, 1:-1] if k == "cmax" else base[k]
        w = w / o.abs().max()
        if k == "cmax":  # levels 0..nAlts+1 only
            w[:, 0] = 0.0
            w[:, -1] = 0.0
        wn[k] = w.contiguous()

    def f(st):
        out = adjoint._prep(solver, st, geometry, pp)
        return sum(float((out[k] * wn[k]).sum()) for k in PREP + ("cp",))

    gprep = {k: wn[k] for k in PREP}
    gcp = wn["cp"]
    gin = {k: torch.zeros_like(d[k]) for k in STATE}
    adjoint.prepare_adjoint(solver, d, gprep, gin, gcp=gcp, pp=pp)
    torch.cuda.synchronize()
    solver.check()
    # directions with random signs: scaling all densities together leaves
    # Gamma and cp unchanged, a direction whose derivative is rounding
    for field in ("nds", "temp", "rho", "vvel"):
        x = state[field]
        v = x * torch.from_numpy(rng.standard_normal(tuple(x.shape))).cuda()
        h = 1e-6
        plus = dict(state)
        minus = dict(state)
        plus[field] = x + h * v
        minus[field] = x - h * v
        fd = (f(plus) - f(minus)) / (2 * h)
        adj = float((gin[field] * v).sum())
        rel = abs(adj - fd) / abs(adj)
        print(field, adj, fd, rel)
        assert rel < 1e-6, (field, adj, fd)


def test_source_adjoint_bits(solver):
    """gitm_vert_source_adjoint against the heating source's adjoint written
    out in numpy in the same order of operations (a = lam * (Dt / (cp *
    TempUnit)) on levels 1..nAlts, g_heat += a, g_tu -= a Q / TempUnit, g_cp
    = -(a Q / cp)): the same bits, also with the prep held (g_heat only).
    Mismatched or partial arguments are refused before any launch."""
    import torch
    from gitm_amd import adjoint
    rng = np.random.default_rng(11)
    n, L, dt = 333, 54, 2.0
    lam, heat, gh0, gtu0 = (rng.standard_normal((n, L)) for _ in range(4))
    cp = 1000.0 + 500.0 * rng.random((n, L))
    tu = 1.0 + 30.0 * rng.random((n, L))
    a = np.zeros((n, L))
    a[:, 2:-2] = lam[:, 2:-2] * (dt / (cp[:, 2:-2] * tu[:, 2:-2]))
    b = a * heat
    want_gh, want_gtu, want_gcp = gh0 + a, gtu0 - b / tu, -(b / cp)
    dv = {k: torch.from_numpy(v.copy()).cuda()
          for k, v in dict(lam=lam, heat=heat, gh=gh0, gtu=gtu0, cp=cp, tu=tu).items()}
    gcp = torch.full_like(dv["cp"], float("nan"))
    adjoint.source_adjoint(solver, dv["lam"], dv["cp"], dv["tu"], dv["heat"], dt, dv["gh"],
                           dv["gtu"], gcp)
    solver.check()
    assert np.array_equal(dv["gh"].cpu().numpy(), want_gh)
    assert np.array_equal(dv["gtu"].cpu().numpy(), want_gtu)
    assert np.array_equal(gcp.cpu().numpy(), want_gcp)
    gh = torch.from_numpy(gh0.copy()).cuda()
    adjoint.source_adjoint(solver, dv["lam"], dv["cp"], dv["tu"], dv["heat"], dt, gh)
    assert np.array_equal(gh.cpu().numpy(), want_gh)
    with pytest.raises(ValueError):
        adjoint.source_adjoint(solver, dv["lam"], dv["cp"], dv["tu"], dv["heat"], dt, gh,
                               dv["gtu"], None)
    lib = adjoint._lib()
    p = ctypes.cast(gh.data_ptr(), ctypes.POINTER(ctypes.c_double))
    assert lib.gitm_vert_source_adjoint(solver._h, n, L - 4, dt, p, p, p, p, p, p, None) != 0
    assert lib.gitm_vert_source_adjoint(solver._h, -1, L - 4, dt, p, p, p, p, p, None, None) != 0
    assert lib.gitm_vert_source_adjoint(solver._h, n, L - 4, dt, None, p, p, p, p, None,
                                        None) != 0


def test_adjoint_record_chunks(solver):
    """A small record budget runs the columns in chunks: the same bits.  A
    column's record, work area and boundary record are (5 x 125 + 54) x 64
    + 300 doubles (350,048 B): 64 MiB holds 191 columns, so the 324 columns
    run as a chunk of 191 and one of 133."""
    import torch
    from gitm_amd import adjoint, vertical
    cols = synth.make_columns(synth.CONFIGS["cpu_324x50"], seed=4)
    p = synth.StepParams()
    d = vertical.to_device(cols)
    go = {k: torch.ones_like(d[k]) / d[k].abs().max() for k in STATE}
    a = {k: v.clone() for k, v in adjoint.adjoint_step(solver, d, go, p).items()}
    adjoint.set_adjoint_memory(solver, 64 << 20)
    try:
        b = adjoint.adjoint_step(solver, d, go, p)
        solver.check()
    finally:
        adjoint.set_adjoint_memory(solver, 0)  # the default again
        adjoint.release_adjoint_memory(solver)
    for k in STATE:
        assert torch.equal(a[k], b[k]), k


@pytest.fixture(scope="module")
def window_case():
    """The differentiation configuration: the full 5-degree grid."""
    import torch
    from gitm_amd import adjoint, vertical
    s = vertical.VerticalSolver()
    cols = synth.make_columns(synth.CONFIGS["5deg_2592x50"], seed=5)
    d = vertical.to_device(cols)
    fixed = {k: d[k] for k in adjoint.GEOMETRY}
    heat = torch.from_numpy(adjoint.heating_profile(cols["alt"])).cuda()
    W = torch.from_numpy(adjoint.level_weights(cols["alt"], 400e3)).cuda()
    p = synth.StepParams()
    nc = W.shape[0]

    def terms(state, q):
        """Each column's term of the loss (mean mass density at 400 km)."""
        out = adjoint.window(state, q, s, fixed, p, 10)
        return (out["rho"] * W).sum(dim=1) / nc

    state = {k: d[k].clone().requires_grad_(True) for k in STATE}
    q = heat.clone().requires_grad_(True)
    J = terms(state, q).sum()
    J.backward()
    grads = {k: state[k].grad.clone() for k in STATE}
    grads["heat"] = q.grad.clone()
    base = {k: d[k].clone() for k in STATE}
    yield s, terms, base, heat, grads
    s.close()


@pytest.mark.parametrize("field", ["temp", "heat"])
def test_window_random_cells(window_case, field):
    """32 random cells, one per column (so that three runs give them all),
    each against its own column's term of the loss."""
    import torch
    s, terms, base, heat, grads = window_case
    x = (heat if field == "heat" else base["temp"]).cpu().numpy()
    g = grads[field].cpu().numpy()
    with torch.no_grad():
        t0 = terms(base, heat).cpu().numpy()
    # the step of each cell: 1e-6 of the temperature; 0.01 W/kg of heating
    # (an absolute step: the heating is ~0 where the loss is sensitive to it)
    step = 1e-6 * x if field == "temp" else np.full_like(x, 1e-2)
    # the change of each cell's column term under its step, relative to the term
    moved = np.abs(g * step) / np.abs(t0)[:, None]
    nc, L = x.shape
    rng = np.random.default_rng(21 if field == "temp" else 22)
    picks = []
    for c in rng.permutation(nc):
        ok = np.nonzero(moved[c, 2:L - 2] > 1e-8)[0] + 2
        if len(ok):
            picks.append((c, int(rng.choice(ok))))
        if len(picks) == 32:
            break
    assert len(picks) == 32
    v = torch.zeros_like(heat)
    for c, k in picks:
        v[c, k] = float(step[c, k])

    def run(sign):
        with torch.no_grad():
            st = {k: t.clone() for k, t in base.items()}
            q = heat.clone()
            if field == "heat":
                q = q + sign * v
            else:
                st["temp"] = st["temp"] + sign * v
            return terms(st, q).cpu().numpy()

    tp, tm = run(1.0), run(-1.0)
    worst, kinks = 0.0, 0
    for c, k in picks:
        adj = g[c, k] * step[c, k]
        central = (tp[c] - tm[c]) / 2.0
        fwd, bwd = tp[c] - t0[c], t0[c] - tm[c]
        err = abs(adj - central) / abs(adj)
        if err >= 1e-6:
            # a switch of the loss (a limiter, upwind choice, the maximum in
            # cMax, the eddy ramp) within the step: the adjoint is the
            # derivative on the smooth side
            kinks += 1
            err = min(abs(adj - fwd), abs(adj - bwd)) / abs(adj)
            assert err < 2e-6, (field, c, k, adj, central, fwd, bwd, err)
        worst = max(worst, err)
    print(field, "32 cells, worst relative error", worst, "cells at a switch", kinks)
    assert kinks <= 4


def test_window_directional_derivatives(window_case):
    """Whole-field directions of the temperature and the densities.  Every
    cell moves, so a kink anywhere in the grid enters the differences: the
    bar is 1e-6 or twice the differences' own spread over h/10..10h."""
    import torch
    s, terms, base, heat, grads = window_case
    for field in ("temp", "nds"):
        v = base[field].clone()

        def fd(h):
            def run(sign):
                with torch.no_grad():
                    st = {k: t.clone() for k, t in base.items()}
                    st[field] = st[field] + sign * h * v
                    return float(terms(st, heat).sum())
            return (run(1.0) - run(-1.0)) / (2.0 * h)

        f1, f2, f3 = fd(1e-6), fd(1e-5), fd(1e-7)
        adj = float((grads[field] * v).sum())
        rel = abs(adj - f1) / abs(adj)
        spread = max(abs(f1 - f2), abs(f1 - f3)) / abs(adj)
        print(field, adj, f1, f2, f3, rel, spread)
        assert rel < max(1e-6, 2.0 * spread), (field, adj, f1, rel, spread)


def test_window_held_prep_matches_cpu_reference():
    """with_prep=False on 4 columns: the GPU window's gradients against the
    same window on the CPU (the oracle forward; backwards, per step, the
    host build of the taped adjoint and the source's adjoint)."""
    import torch
    from gitm_amd import adjoint, vertical
    from oracle import cpu
    from tests.emu import emu
    s = vertical.VerticalSolver()
    cols = subset(synth.make_columns(synth.CONFIGS["5deg_2592x50"], seed=5), [3, 700, 1500, 2591])
    d = vertical.to_device(cols)
    for k in ("press", "cp"):
        d[k] = torch.zeros_like(d["temp"])
    s.prepare(d, synth.PrepParams())
    fixed = {k: d[k] for k in adjoint.GEOMETRY + PREP + ("cp",)}
    heat_np = adjoint.heating_profile(cols["alt"])
    heat = torch.from_numpy(heat_np).cuda()
    W = adjoint.level_weights(cols["alt"], 400e3)
    p = synth.StepParams()
    n = 10
    state = {k: d[k].clone().requires_grad_(True) for k in STATE}
    q = heat.clone().requires_grad_(True)
    out = adjoint.window(state, q, s, fixed, p, n, with_prep=False)
    J = (out["rho"] * torch.from_numpy(W).cuda()).sum()
    J.backward()
    # the CPU window
    host = {k: d[k].cpu().numpy() for k in adjoint.GEOMETRY + PREP + ("cp",)}
    cp_tu = host["cp"] * host["tempunit"]
    st = {k: np.array(cols[k]) for k in STATE}
    cks = []
    for _ in range(n):
        cks.append({k: v.copy() for k, v in st.items()})
        o = cpu.advance(dict(st, **host), p, solver=1)
        st = {k: o[k] for k in STATE}
        st["temp"] = st["temp"].copy()
        st["temp"][:, 2:-2] += p.dt * heat_np[:, 2:-2] / cp_tu[:, 2:-2]
    lam = {k: np.zeros_like(st[k]) for k in STATE}
    lam["rho"] = W.copy()
    gq = np.zeros_like(heat_np)
    for ck in reversed(cks):
        gq[:, 2:-2] += lam["temp"][:, 2:-2] * p.dt / cp_tu[:, 2:-2]
        lam, _, _ = emu.adjoint(dict(ck, **host), p, lam)
    for k in STATE:
        err = _normwise(state[k].grad.cpu().numpy(), lam[k])
        print(k, err)
        assert err < 1e-9, (k, err)
    err = _normwise(q.grad.cpu().numpy(), gq)
    print("heat", err)
    assert err < 1e-9, ("heat", err)
    s.close()
