"""The oracle's sources step (ov_sources_step, ov_conduction_1d), pinned to
the reference build and by what the mathematics fixes.

* Bit for bit against the reference's calc_thermal_conduction and
  calc_viscosity (src/calc_conduction.f90, src/calc_viscosity.f90, compiled
  from the reference sources: oracle/_ref/gitm_sources_ref): Conduction,
  Viscosity and VerticalViscosityS, laminar and turbulent, on the golden
  fixtures tests/golden/sources_*.npz and on fresh columns when the
  reference build is present.

* calc_conduction_1d (calc_conduction.f90:132-244) on a uniform grid with a
  constant radius reduces to the textbook implicit step of
  MulFac dT/dt = D T'' + D' T' (3-point centred differences, T fixed at the
  bottom, a mirror ghost at the top for the zero gradient): solved here with
  numpy from that textbook form.
* On stretched grids with the spherical r^2 factors a constant profile stays
  constant (the rows' weights cancel); the steady state T = A + B/r of
  constant D is held to rounding on uniform grids and converges at second
  order on stretched ones.
* The two-stage combination (2H - F) - Q of calc_thermal_conduction and
  calc_viscosity is second order in Dt: against many small implicit steps
  the error falls by four when Dt halves (one implicit step alone: by two).
* The updates of add_sources: Temperature(1:nAlts) + Dt * heating, then
  every level + Conduction; the winds + Viscosity; Velocity(up) + what
  calc_viscosity leaves in Viscosity(up), the last species' full-step solve.
"""
from __future__ import annotations

import glob
import json
import os

import numpy as np
import pytest

from gitm_amd import synth
from oracle import cpu, ref
from tests.helpers import subset

GOLDEN = sorted(glob.glob(os.path.join(os.path.dirname(__file__), "golden", "sources_*.npz")))
INCREMENTS = ("conduction", "viscosity", "vert_viscosity")


def test_sources_fixtures_present():
    assert len(GOLDEN) >= 2


@pytest.mark.parametrize("path", GOLDEN, ids=[os.path.basename(p)[:-4] for p in GOLDEN])
def test_oracle_bitwise_equals_reference_sources(path):
    z = np.load(path)
    meta = json.loads(str(z["meta"]))
    cols = {k[3:]: z[k] for k in z.files if k.startswith("in_")}
    want = {k[4:]: z[k] for k in z.files if k.startswith("out_")}
    got = cpu.sources(cols, meta["dt"], synth.species_masses(),
                      use_turbulent_cond=meta["use_turbulent_cond"], apply=False)
    L = cols["temp"].shape[1]
    for k in INCREMENTS:  # levels 0..nAlts+1
        assert np.array_equal(got[k][..., 1:L - 1], want[k][..., 1:L - 1]), k


@pytest.mark.skipif(not ref.sources_available(), reason="reference build (oracle/_ref) absent")
@pytest.mark.parametrize("turbulent", [False, True])
def test_oracle_bitwise_equals_reference_sources_fresh(turbulent):
    """Four blocks (324 columns), the reference's own ViscCoef and ViscCoefS
    (calc_viscosity_coef) fed to the oracle."""
    cols = synth.make_columns(synth.CONFIGS["cpu_324x50"], seed=33 + int(turbulent))
    cols.update(synth.source_fields(cols, seed=34))
    want = ref.sources(cols, 2.0, turbulent)
    cols["visc_coef"], cols["visc_coef_s"] = want["visc_coef"], want["visc_coef_s"]
    got = cpu.sources(cols, 2.0, synth.species_masses(), use_turbulent_cond=turbulent,
                      apply=False)
    L = cols["temp"].shape[1]
    for k in INCREMENTS:
        assert np.array_equal(got[k][..., 1:L - 1], want[k][..., 1:L - 1]), k


def textbook(nA, dt, q, D, M, h):
    """Implicit step of M (T - q)/dt = D T'' + D' T' on a uniform grid,
    levels 0..nA+1 (T(0) = q(0), zero gradient at nA+1 by a mirror ghost)."""
    n = nA + 2
    A = np.zeros((n, n))
    rhs = np.zeros(n)
    A[0, 0] = 1.0
    rhs[0] = q[0]
    for i in range(1, n):
        lam = dt / M[i]
        if i < n - 1:
            dD = (D[i + 1] - D[i - 1]) / (2 * h)
            A[i, i - 1] = -lam * (D[i] / h ** 2 - dD / (2 * h))
            A[i, i + 1] = -lam * (D[i] / h ** 2 + dD / (2 * h))
            A[i, i] = 1 + 2 * lam * D[i] / h ** 2
        else:  # T(n) = T(n-2) mirror: D (2 T(n-2) - 2 T(n-1)) / h^2
            A[i, i - 1] = -2 * lam * D[i] / h ** 2
            A[i, i] = 1 + 2 * lam * D[i] / h ** 2
        rhs[i] = q[i]
    return np.linalg.solve(A, rhs)


@pytest.mark.parametrize("varD", [False, True])
def test_uniform_grid_is_the_textbook_scheme(varD):
    nA, h = 30, 2000.0
    L = nA + 4
    alt = 100e3 + h * (np.arange(L) - 1.0)
    rd = np.full(L, 1.0e12)  # constant radius: r^2 cancels
    rng = np.random.default_rng(1)
    q = 500.0 + 50.0 * rng.standard_normal(L)
    D = 2.0 + (0.5 * np.sin(np.arange(L) / 3.0) if varD else 0.0)
    M = 1.0 + 0.1 * rng.random(L)
    dt = 3.0e5
    res = cpu.conduction_1d(nA, dt, True, q, np.broadcast_to(D, L), M, alt, rd)
    want = textbook(nA, dt, q[1:nA + 3], np.broadcast_to(D, L)[1:nA + 3], M[1:nA + 3], h)
    got = q[1:nA + 3] + res[1:nA + 3]
    np.testing.assert_allclose(got, want, rtol=1e-12, atol=0)
    assert np.max(np.abs(res)) > 1.0  # the step does move the profile


def test_constant_profile_stays_on_stretched_spherical_grids():
    spec = synth.GridSpec(1, 1, 60, stretched=True)
    cols = subset(synth.make_columns(spec, seed=2), [0])
    f = synth.source_fields(cols)
    L = cols["temp"].shape[1]
    q = np.full(L, 731.25)
    res = cpu.conduction_1d(60, 60.0, True, q, f["kappa_temp"][0] * 1e4,
                            cols["rho"][0] * f["cp"][0], cols["alt"][0], f["raddist"][0])
    assert np.max(np.abs(res[1:L - 1])) <= 1e-12 * 731.25


def test_spherical_steady_state():
    """T = A + B/r solves (r^2 D T')' = 0 for constant D: the scheme holds it
    to rounding on uniform grids, and converges to it at second order on
    geometrically stretched ones (spacing halved, grading kept)."""
    nA = 40
    L = nA + 4
    alt = 100e3 + 5e3 * (np.arange(L) - 1.0)
    rd = 6372e3 + alt
    q = 900.0 + 3.0e9 / rd
    res = cpu.conduction_1d(nA, 1.0e6, False, q, np.full(L, 0.5), np.full(L, 1e-6), alt, rd)
    assert np.max(np.abs(res[2:L - 2])) <= 1e-13 * 1400.0
    errs = []
    for k, nA in enumerate((30, 60, 120, 240)):
        L = nA + 4
        s = 1.04 ** (1.0 / 2 ** k)
        dz = 5e3 / 2 ** k * s ** np.arange(L)
        alt = 100e3 + np.concatenate([[-dz[0]], np.cumsum(dz[:-1])])
        rd = 6372e3 + alt
        q = 900.0 + 3.0e9 / rd
        res = cpu.conduction_1d(nA, 1.0e6, False, q, np.full(L, 0.5), np.full(L, 1e-6),
                                alt, rd)
        errs.append(np.max(np.abs(res[2:L - 2])))
    ratios = [errs[i] / errs[i + 1] for i in range(3)]
    assert all(3.5 < r < 5.0 for r in ratios), (errs, ratios)


def _column():
    cols = subset(synth.make_columns(synth.CONFIGS["cpu_324x50"], seed=4), [7])
    f = synth.source_fields(cols, seed=4)
    cols.update(f)
    return cols


def test_two_stage_combination_is_second_order_in_time():
    """A smooth profile under constant coefficients, changed by ~1 K over
    Dt: the combination's error against a converged reference falls by ~4
    when Dt halves, the single implicit step's by ~2."""
    cols = subset(synth.make_columns(synth.CONFIGS["cpu_324x50"], seed=4), [7])
    alt = cols["alt"][0]
    rd = synth.R_EARTH + alt
    nA = cols["temp"].shape[1] - 4
    T0 = 1000.0 - 800.0 * np.exp(-(alt - 100e3) / 50e3)
    kap = np.full_like(alt, 0.05)
    mul = np.full_like(alt, 1e-6)

    def implicit(q, dt):
        return q + cpu.conduction_1d(nA, dt, True, q, kap, mul, alt, rd)

    def combined(q, dt):
        return 2.0 * implicit(implicit(q, dt / 2), dt / 2) - implicit(q, dt)

    Dt = 100.0
    fine = T0.copy()
    for _ in range(512):
        fine = combined(fine, Dt / 512)
    e2, e1 = [], []
    for n in (1, 2, 4):
        y = T0.copy()
        z = T0.copy()
        for _ in range(n):
            y = combined(y, Dt / n)
            z = implicit(z, Dt / n)
        e2.append(np.max(np.abs((y - fine)[1:nA + 3])))
        e1.append(np.max(np.abs((z - fine)[1:nA + 3])))
    assert np.max(np.abs(fine - T0)) > 0.5
    assert all(3.3 < e2[i] / e2[i + 1] < 5.0 for i in range(2)), e2
    assert all(1.7 < e1[i] / e1[i + 1] < 2.3 for i in range(2)), e1


@pytest.mark.parametrize("turbulent", [False, True])
def test_sources_step_updates(turbulent):
    cols = _column()
    mass = synth.species_masses()
    dt = 2.0
    out = cpu.sources(cols, dt, mass, use_turbulent_cond=turbulent, apply=True)
    ref = cpu.sources(cols, dt, mass, use_turbulent_cond=turbulent, apply=False)
    L = cols["temp"].shape[1]
    nA = L - 4
    # the increments do not depend on apply; the state is untouched without it
    for k in ("conduction", "viscosity", "vert_viscosity"):
        assert np.array_equal(out[k], ref[k]), k
    assert np.array_equal(ref["temp"], cols["temp"])
    # Temperature: heating on 1..nAlts, then the conduction on 0..nAlts+1
    t = cols["temp"][0].copy()
    t[2:nA + 2] = t[2:nA + 2] + dt * cols["heating"][0, 2:nA + 2]
    t[1:nA + 3] = t[1:nA + 3] + out["conduction"][0, 1:nA + 3]
    assert np.array_equal(out["temp"][0], t)
    assert np.array_equal(out["vel"][0, :, 1:nA + 3],
                          cols["vel"][0, :, 1:nA + 3] + out["viscosity"][0, :, 1:nA + 3])
    assert np.array_equal(out["vvel"][0, :, 1:nA + 3],
                          cols["vvel"][0, :, 1:nA + 3] + out["vert_viscosity"][0, :, 1:nA + 3])
    # the turbulent conduction changes the temperature increments only
    other = cpu.sources(cols, dt, mass, use_turbulent_cond=not turbulent, apply=False)
    assert not np.array_equal(other["conduction"], out["conduction"])
    assert np.array_equal(other["viscosity"], out["viscosity"])
