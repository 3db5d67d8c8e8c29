"""The adjoint of the AUSM step on the CPU: the host build of the taped
column code (gitm_amd/csrc/vert_ad.h through tests/emu) against central
finite differences of the oracle, which is an independent restatement of the
reference step; and the host build of the hand-written adjoint
(gitm_amd/csrc/vert_adj.h) against the taped one, its prep inputs'
adjoints against central differences of the oracle.

A loss J = sum of w . (state after one step) is differentiated with respect
to single input cells.  Central differences of the oracle at two step sizes
bound their own error (truncation and the rounding of the perturbed run):
the adjoint must agree with the finer difference to within 1e-6 relative or
to within twice the two differences' disagreement, whichever is larger.
Cells are drawn where the loss is sensitive (|x dJ/dx| above 1e-6 of the
largest one), so that the differences resolve the derivative.
"""
from __future__ import annotations

import numpy as np
import pytest

from gitm_amd import synth
from gitm_amd.adjoint import level_weights, heating_profile
from oracle import cpu
from tests.helpers import subset

STATE = ("nds", "vvel", "vel", "temp", "rho")


@pytest.fixture(scope="module")
def case():
    from tests.emu import emu
    cols = subset(synth.make_columns(synth.CONFIGS["cpu_324x50"], seed=3), [17])
    p = synth.StepParams()
    rng = np.random.default_rng(7)
    # a loss touching every output field: rho near 400 km plus random weights
    w = {k: rng.standard_normal(cols[k].shape) for k in STATE}
    scale = {k: 1.0 / np.max(np.abs(cols[k])) for k in STATE}
    for k in STATE:
        w[k] *= 1e-3 * scale[k]
    w["rho"][0, 31] += 0.5 * scale["rho"]
    w["rho"][0, 32] += 0.5 * scale["rho"]
    gin, nodes, phases = emu.adjoint(cols, p, w)
    return cols, p, w, gin, nodes, phases


def loss(cols, p, w):
    out = cpu.advance(cols, p, solver=1, threads=1)
    return sum(float(np.sum(w[k] * out[k])) for k in STATE)


def central(cols, p, w, field, idx, rel):
    x0 = cols[field][idx]
    h = rel * abs(x0) if x0 != 0 else rel
    a = {k: np.array(v) for k, v in cols.items()}
    a[field][idx] = x0 + h
    b = {k: np.array(v) for k, v in cols.items()}
    b[field][idx] = x0 - h
    return (loss(a, p, w) - loss(b, p, w)) / (2 * h)


def test_tape_size(case):
    _, _, _, _, nodes, phases = case
    # one step of a 54-level column: about 10k nodes and 85 barriers per lane
    assert 5000 < nodes < 12288 + 8 * 50
    assert 40 < phases < 128


@pytest.mark.parametrize("field", STATE)
def test_adjoint_matches_oracle_differences(case, field):
    cols, p, w, gin, _, _ = case
    g = gin[field]
    sens = np.abs(g * cols[field])
    cand = np.argwhere(sens > 1e-6 * sens.max())
    # a fixed seed per field (str hashes change from process to process)
    rng = np.random.default_rng(sum(map(ord, field)))
    picks = cand[rng.choice(len(cand), size=min(4, len(cand)), replace=False)]
    for idx in map(tuple, picks):
        f1 = central(cols, p, w, field, idx, 1e-6)
        f2 = central(cols, p, w, field, idx, 1e-5)
        err = abs(g[idx] - f1) / abs(g[idx])
        bound = max(1e-6, 2.0 * abs(f1 - f2) / abs(g[idx]))
        assert err <= bound, (field, idx, g[idx], f1, f2, err, bound)


def test_adjoint_is_linear_in_the_seed(case):
    """The adjoint map is linear: seeds w and 3w give 3x the adjoints (to
    rounding: the order of the atomic adds varies from run to run), and a
    zero seed gives zero."""
    from tests.emu import emu
    cols, p, w, gin, _, _ = case
    g3, _, _ = emu.adjoint(cols, p, {k: 3.0 * w[k] for k in STATE})
    for k in STATE:
        np.testing.assert_allclose(g3[k], 3.0 * gin[k], rtol=1e-12,
                                   atol=1e-14 * float(np.max(np.abs(3.0 * gin[k]))))
    g0, _, _ = emu.adjoint(cols, p, {k: np.zeros_like(w[k]) for k in STATE})
    assert all(np.all(g0[k] == 0.0) for k in STATE)


def test_level_weights_and_heating_profile():
    alt = np.array([[90e3, 100e3, 110e3, 395e3, 405e3, 415e3, 500e3, 510e3]])
    w = level_weights(alt, 400e3)
    assert w[0, 3] == pytest.approx(0.5) and w[0, 4] == pytest.approx(0.5)
    assert w.sum() == pytest.approx(1.0)
    q = heating_profile(np.array([150e3, 170e3]))
    assert q[0] == pytest.approx(1e-5) and q[1] == pytest.approx(1e-5 * np.exp(-1.0))


def test_source_adjoint_argument_checks():
    """The heating source's adjoint takes contiguous float64 device tensors
    of one shape, with the prep's cp and TempUnit adjoints both or neither:
    anything else is refused on the host, before the library is reached."""
    import torch
    from gitm_amd import adjoint
    t = torch.zeros(3, 8, dtype=torch.float64)
    with pytest.raises(ValueError):
        adjoint.source_adjoint(None, t, t, t, t, 1.0, t)  # host tensors
    with pytest.raises(ValueError, match="go together"):
        adjoint.source_adjoint(None, t, t, t, t, 1.0, t, t, None)


# ---------------------------------------------------------------------------
# The hand-written adjoint (gitm_amd/csrc/vert_adj.h), host build.
# ---------------------------------------------------------------------------
@pytest.mark.parametrize("device_math", [False, True])
@pytest.mark.parametrize("n_alts,seed", [(50, 3), (20, 4), (100, 5)])
def test_hand_adjoint_matches_taped_build(n_alts, seed, device_math):
    """The hand-written adjoint against the taped one (both host builds) for
    a random seed on every output: 1e-12 normwise per field.  100 levels use
    two wavefronts; the device-math forms (log-space sums, friction powers as
    exp(e ln T)) give the same derivatives to rounding."""
    from tests.emu import emu
    cols = synth.make_columns(synth.GridSpec(1, 1, n_alts), seed=seed)
    cols = subset(cols, [0, 40] if n_alts != 100 else [7])
    p = synth.StepParams()
    rng = np.random.default_rng(seed)
    gout = {k: rng.standard_normal(cols[k].shape) / np.max(np.abs(cols[k])) for k in STATE}
    ref, _, _ = emu.adjoint(cols, p, gout)
    hand = emu.adjoint_hand(cols, p, gout, device_math=device_math)
    for k in STATE:
        err = np.max(np.abs(hand[k] - ref[k])) / np.max(np.abs(ref[k]))
        assert err < 1e-12, (k, err)


@pytest.mark.parametrize("flags", [dict(use_neutral_friction=False),
                                   dict(use_coriolis=False),
                                   dict(dt=5.0)])
def test_hand_adjoint_options(flags):
    from tests.emu import emu
    cols = subset(synth.make_columns(synth.GridSpec(1, 1, 30), seed=9), [3])
    p = synth.StepParams(**flags)
    rng = np.random.default_rng(2)
    gout = {k: rng.standard_normal(cols[k].shape) / np.max(np.abs(cols[k])) for k in STATE}
    ref, _, _ = emu.adjoint(cols, p, gout)
    hand = emu.adjoint_hand(cols, p, gout)
    for k in STATE:
        err = np.max(np.abs(hand[k] - ref[k])) / np.max(np.abs(ref[k]))
        assert err < 1e-12, (flags, k, err)


@pytest.fixture(scope="module")
def prep_case():
    from tests.emu import emu
    cols = subset(synth.make_columns(synth.CONFIGS["cpu_324x50"], seed=3), [0])
    p = synth.StepParams()
    rng = np.random.default_rng(2)
    w = {k: rng.standard_normal(cols[k].shape) / np.max(np.abs(cols[k])) for k in STATE}
    _, gf = emu.adjoint_hand(cols, p, w, fixed=True)
    return cols, p, w, gf


# (field, level): every prep input at levels where it enters -- the bottom
# ghost levels (MeanMajorMass of the level -1 fill and the hydrostatic
# background; Gamma's face values), interior levels and the top
@pytest.mark.parametrize("field,lev", [("tempunit", 1), ("tempunit", 30), ("tempunit", 53),
                                       ("gamma", 0), ("gamma", 2), ("gamma", 25),
                                       ("gamma", 52), ("mmm", 0), ("mmm", 1), ("cmax", 2),
                                       ("cmax", 3), ("cmax", 20), ("eddy", 2), ("eddy", 20)])
def test_prep_input_adjoints_match_oracle_differences(prep_case, field, lev):
    """The adjoints of the column-prep inputs (TempUnit, Gamma, MeanMajorMass,
    cMax, the eddy coefficient) against central differences of the oracle
    step.  Zero inputs (the eddy coefficient above its pressure ramp) are
    stepped by an absolute h."""
    cols, p, w, gf = prep_case
    g = gf[field][0, lev]
    x0 = cols[field][0, lev]
    f = {}
    for rel in (1e-6, 1e-5):
        h = rel * abs(x0) if x0 != 0 else 1e3 * rel
        a = {k: np.array(v) for k, v in cols.items()}
        a[field][0, lev] = x0 + h
        b = {k: np.array(v) for k, v in cols.items()}
        b[field][0, lev] = x0 - h
        f[rel] = (loss(a, p, w) - loss(b, p, w)) / (2 * h)
    err = abs(g - f[1e-6]) / abs(g)
    bound = max(1e-6, 2.0 * abs(f[1e-6] - f[1e-5]) / abs(g))
    assert err <= bound, (field, lev, g, f, err, bound)
