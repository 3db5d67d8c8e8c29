"""GPU adjoint of the AUSM step (gitm_vert_adjoint) and the differentiable
window (gitm_amd.adjoint.VerticalWindow).

* One step against the CPU FP64 reference gradient: the host build of the
  same taped column code (tests/emu), itself checked against finite
  differences of the oracle in test_adjoint_cpu.py.
* Directional derivatives of a 10-step window (AUSM step + heating source,
  the loss the mean mass density at 400 km) along whole-field directions of
  the temperature and of the heating rate, against central differences:
  1e-6 relative.  Along a direction that moves every cell the change of the
  loss is large against the rounding of the perturbed runs, which a single
  cell's is not (a cell's share of the loss is too small for a 1e-6
  difference to resolve it in FP64).
* Single cells, the most sensitive ones: the adjoint within the differences'
  own error (two step sizes), or 1e-6.
"""
from __future__ import annotations

import numpy as np
import pytest

from gitm_amd import synth
from tests.helpers import subset

pytestmark = pytest.mark.gpu

STATE = ("nds", "vvel", "vel", "temp", "rho")


def test_adjoint_step_matches_cpu_taped_build(solver):
    import torch
    from gitm_amd import adjoint, vertical
    from tests.emu import emu
    cols = subset(synth.make_columns(synth.CONFIGS["cpu_324x50"], seed=3), [0, 17, 200])
    p = synth.StepParams()
    rng = np.random.default_rng(11)
    gout = {k: rng.standard_normal(cols[k].shape) / np.max(np.abs(cols[k])) for k in STATE}
    ref, _, _ = emu.adjoint(cols, p, gout)
    d = vertical.to_device(cols)
    go = {k: torch.from_numpy(gout[k]).cuda() for k in STATE}
    gi = adjoint.adjoint_step(solver, d, go, p)
    for k in STATE:
        a, b = gi[k].cpu().numpy(), ref[k]
        err = np.max(np.abs(a - b)) / np.max(np.abs(b))
        assert err < 1e-12, (k, err)
    # in place (gin = gout) gives the same
    gi2 = adjoint.adjoint_step(solver, d, go, p, gin=go)
    for k in STATE:
        np.testing.assert_allclose(gi2[k].cpu().numpy(), gi[k].cpu().numpy(), rtol=1e-13,
                                   atol=1e-13 * float(gi[k].abs().max()))


def test_adjoint_chunks(solver):
    """A small tape budget runs the columns in many chunks: same adjoints.
    One column's tape takes 12,688 nodes x 32 B plus the phase counts, x 54
    lanes (21.95 MB): 128 MiB holds 6 columns (36 = 6 x 6), 112 MiB holds 5,
    which leaves a last chunk of one column, whose adjoints must start at
    zero over the whole tape again."""
    import torch
    from gitm_amd import adjoint, vertical
    cols = subset(synth.make_columns(synth.CONFIGS["cpu_324x50"], seed=4), np.arange(0, 324, 9))
    p = synth.StepParams()
    d = vertical.to_device(cols)
    go = {k: torch.ones_like(d[k]) / d[k].abs().max() for k in STATE}
    a = adjoint.adjoint_step(solver, d, go, p)
    for budget in (128 << 20, 112 << 20):
        adjoint.set_adjoint_memory(solver, budget)
        try:
            b = adjoint.adjoint_step(solver, d, go, p)
        finally:
            adjoint.set_adjoint_memory(solver, 0)  # the default again
        for k in STATE:
            scale = float(a[k].abs().max())
            assert float((a[k] - b[k]).abs().max()) <= 1e-13 * scale, (budget, k)


@pytest.fixture(scope="module")
def window_case():
    import torch
    from gitm_amd import adjoint, vertical
    s = vertical.VerticalSolver()
    cols = subset(synth.make_columns(synth.CONFIGS["5deg_2592x50"], seed=5),
                  np.arange(0, 2592, 9))
    d = vertical.to_device(cols)
    for k in ("press", "cp"):
        d[k] = torch.zeros_like(d["temp"])
    s.prepare(d, synth.PrepParams())
    fixed = {k: d[k] for k in adjoint.FIXED}
    fixed["cp"] = d["cp"]
    heat = torch.from_numpy(adjoint.heating_profile(cols["alt"])).cuda()
    W = torch.from_numpy(adjoint.level_weights(cols["alt"], 400e3)).cuda()
    p = synth.StepParams()

    def loss(state, q):
        out = adjoint.window(state, q, s, fixed, p, 10)
        return (out["rho"] * W).sum() / W.shape[0]

    state = {k: d[k].clone().requires_grad_(True) for k in STATE}
    q = heat.clone().requires_grad_(True)
    J = loss(state, q)
    J.backward()
    grads = {k: state[k].grad.clone() for k in STATE}
    grads["heat"] = q.grad.clone()
    base = {k: d[k].clone() for k in STATE}
    yield s, loss, base, heat, float(J.detach()), grads
    s.close()


def central(loss, base, heat, field, v, h):
    import torch
    with torch.no_grad():
        def shifted(sign):
            st = {k: t.clone() for k, t in base.items()}
            q = heat.clone()
            if field == "heat":
                q = q + sign * h * v
            else:
                st[field] = st[field] + sign * h * v
            return float(loss(st, q))
        return (shifted(1.0) - shifted(-1.0)) / (2.0 * h)


@pytest.mark.parametrize("field,band", [("temp", None), ("temp", (20, 40)), ("nds", None),
                                        ("heat", None)])
def test_window_directional_derivatives(window_case, field, band):
    import torch
    s, loss, base, heat, J, grads = window_case
    rng = np.random.default_rng(3)
    x = heat if field == "heat" else base[field]
    v = x.clone()
    if band is not None:  # random signs on a band of levels
        mask = torch.zeros_like(v)
        mask[..., band[0]:band[1]] = torch.from_numpy(
            rng.choice([-1.0, 1.0], size=mask[..., band[0]:band[1]].shape)).cuda()
        v = v * mask
    # the heating moves the temperature by ~1e-8 K per step: a relative step
    # of 1e3 of the profile keeps the response linear and well above rounding
    h = 1e3 if field == "heat" else 1e-6
    fd = central(loss, base, heat, field, v, h)
    fd2 = central(loss, base, heat, field, v, 10.0 * h)
    fd3 = central(loss, base, heat, field, v, 0.1 * h)
    adj = float((grads[field] * v).sum())
    rel = abs(adj - fd) / abs(adj)
    spread = max(abs(fd - fd2), abs(fd - fd3)) / abs(adj)
    print(field, band, "adjoint", adj, "fd", fd, fd2, fd3, "rel", rel, "fd spread", spread,
          "J", J)
    if field != "heat":  # (the heating moves the loss by ~1e-14 of itself)
        assert abs(adj) > 1e-6 * abs(J)  # the direction does move the loss
    assert rel < max(1e-6, 2.0 * spread), (adj, fd, rel, spread)


def test_window_single_cells(window_case):
    s, loss, base, heat, J, grads = window_case
    import torch
    g = grads["temp"].cpu().numpy()
    x = base["temp"].cpu().numpy()
    sens = np.abs(g * x)
    order = np.argsort(sens.ravel())[::-1][:200]
    rng = np.random.default_rng(5)
    picks = rng.choice(order, size=8, replace=False)
    worst = 0.0
    for flat in picks:
        c, k = np.unravel_index(flat, g.shape)
        v = torch.zeros_like(base["temp"])
        v[c, k] = float(x[c, k])
        f1 = central(loss, base, heat, "temp", v, 1e-6)
        f2 = central(loss, base, heat, "temp", v, 1e-5)
        a = g[c, k] * x[c, k]
        err = abs(a - f1) / abs(a)
        bound = max(1e-6, 2.0 * abs(f1 - f2) / abs(a))
        worst = max(worst, err / bound)
        assert err <= bound, (c, k, a, f1, f2, err, bound)
    print("worst err/bound", worst)
