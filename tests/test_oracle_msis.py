"""UseMSIS lower boundary (set_vertical_bcs.Earth.f90:113-137): the oracle's
branch pinned to the reference build, by what the reference's code fixes,
and the host build of the kernel against it bit for bit.

Against the reference: the reference build runs the branch itself, calling
its own msis_bcs (init_msis.Earth.f90:367, MSIS-00 GTD7 with no HWM winds)
at every boundary pass; the oracle, given what msis_bcs returns for each
column's level 0, reproduces the reference's step bit for bit (fresh columns
here, and the golden fixtures ausm_msis_3step and rusanov_msis_gswm_2step).
That also pins what the invariants below cannot: the hydrostatic fill of
level -1 uses the entry NS(0), not the MSIS value (:93, :198-223).

The reference calls msis_bcs for levels -1 and 0 at every boundary pass; the
values depend on time, place and solar indices only, so the caller computes
them (gitm_amd.synth.msis_bc_values stands in for MSIS-90 here).  What the
branch must then do, and these tests check on the oracle's output:
* LogNS(0) of the non-photochemical species, Temp(0) and the east and north
  winds of level 0 are the MSIS values after the step (the last boundary pass
  sets them and nothing after it changes level 0);
* LogNS(-1) of O equals the MSIS LogNS(0) of O (zero gradient, :219-222);
* Vel(-1) east and north is the 4th-order one-sided extrapolation from levels
  0..4 (:364-375), with the oracle's one-sided weights (themselves exact for
  quartics, test_oracle_math.py);
* given MSIS values equal to the column's own level-0 state and no
  horizontal wind, the branch changes nothing against UseMSIS = F.
"""
from __future__ import annotations

import ctypes
import dataclasses

import numpy as np
import pytest

from gitm_amd import synth
from oracle import cpu, ref
from tests.helpers import subset


@pytest.mark.skipif(not ref.available(), reason="reference build (oracle/_ref) absent")
@pytest.mark.parametrize("solver", [0, 1])
def test_oracle_bitwise_equals_reference_msis(solver):
    cols = subset(synth.make_columns(synth.CONFIGS["cpu_324x50"], seed=61 + solver),
                  np.arange(3, 324, 17))
    nc = cols["temp"].shape[0]
    rng = np.random.default_rng(solver)
    cols["lat"] = np.arccos(np.clip(cols["coslat"], -1.0, 1.0)) * \
        np.where(rng.random(nc) < 0.5, -1.0, 1.0)
    cols["lon"] = rng.uniform(0.0, 2.0 * np.pi, nc)
    env = dict(julian_day=300, utime=72000.0, f107=120.0, f107a=130.0)
    p = synth.StepParams(use_msis=True)
    want = ref.advance(cols, p, solver=solver, n_steps=3, msis_env=env)
    cols["msis"] = want.pop("msis")
    got = cpu.advance(cols, p, solver=solver, n_steps=3)
    for k in want:
        assert np.array_equal(got[k], want[k]), k

STATE = ("nds", "vvel", "vel", "temp", "rho")


def _case(seed=31, ncols=12, nalts=50, stretched=False):
    spec = synth.GridSpec(1, 1, nalts, stretched=stretched)
    cols = subset(synth.make_columns(spec, seed=seed), np.arange(ncols))
    cols["msis"] = synth.msis_bc_values(cols, seed=seed)
    return cols


@pytest.mark.parametrize("solver", [1, 0])
def test_level_zero_takes_the_msis_values(solver):
    cols = _case()
    p = dataclasses.replace(synth.StepParams(), use_msis=True)
    out = cpu.advance(cols, p, solver=solver)
    m = cols["msis"]
    for sp in (0, 1, 2, 5):  # non-photochemical
        assert np.array_equal(out["logns"][:, sp, 1], m[:, sp]), sp
    assert np.array_equal(out["logns"][:, 0, 0], m[:, 0])  # O below: zero gradient
    np.testing.assert_allclose(out["temp"][:, 1] * cols["tempunit"][:, 1], m[:, 6],
                               rtol=2.3e-16, atol=0)
    assert np.array_equal(out["vel"][:, 0, 1], m[:, 7])
    assert np.array_equal(out["vel"][:, 1, 1], m[:, 8])
    # the photochemical species are extrapolated from above, not MSIS's
    assert not np.array_equal(out["logns"][:, 3, 1], m[:, 3])


def test_ghost_winds_are_extrapolated():
    cols = _case(stretched=True, nalts=60)
    p = dataclasses.replace(synth.StepParams(), use_msis=True)
    out = cpu.advance(cols, p, solver=1)
    lib = cpu.lib()
    lib.ov_bc_coefs.argtypes = [ctypes.c_double] * 4 + [ctypes.c_int,
                                                        ctypes.POINTER(ctypes.c_double)]
    for c in range(cols["temp"].shape[0]):
        alt = cols["alt"][c]
        dF = np.diff(alt)  # dF[i] = Alt(level i) - Alt(level i - 1), i = level + 1
        C = (ctypes.c_double * 5)()
        lib.ov_bc_coefs(dF[1], dF[2], dF[3], dF[4], 0, C)
        for d in range(2):
            v = out["vel"][c, d]
            dv = v[1] * C[0] + v[2] * C[1] + v[3] * C[2] + v[4] * C[3] + v[5] * C[4]
            assert v[0] == v[1] - dF[0] * dv
        assert np.any(out["vel"][c, :2, 0] != 0.0)


def test_own_values_change_nothing():
    cols = _case(seed=33)
    cols["vel"][:, :2, :] = 0.0  # no horizontal wind
    cols["msis"][:, :6] = np.log(cols["nds"][:, :, 1])
    cols["msis"][:, 6] = cols["temp"][:, 1] * cols["tempunit"][:, 1]
    cols["msis"][:, 7:] = 0.0
    p = dataclasses.replace(synth.StepParams(), use_coriolis=False)
    plain = cpu.advance(cols, p, solver=1, n_steps=3)
    with_msis = cpu.advance(cols, dataclasses.replace(p, use_msis=True), solver=1, n_steps=3)
    for k in STATE:
        assert np.array_equal(plain[k], with_msis[k]), k
    # a different MSIS temperature does change the column
    cols["msis"][:, 6] *= 1.01
    moved = cpu.advance(cols, dataclasses.replace(p, use_msis=True), solver=1, n_steps=3)
    assert not np.array_equal(moved["temp"], plain["temp"])


@pytest.mark.parametrize("solver", [1, 0])
def test_emulated_kernel_matches_oracle_with_msis(solver):
    from tests.emu import emu
    cols = _case(seed=35, ncols=6, nalts=100, stretched=True)
    p = dataclasses.replace(synth.StepParams(), use_msis=True)
    out = emu.advance(cols, p, solver=solver, n_steps=3)
    ref = cpu.advance(cols, p, solver=solver, n_steps=3)
    for k in STATE:
        assert np.array_equal(out[k], ref[k]), k


def test_msis_needs_its_values():
    from gitm_amd import _abi
    from tests.emu import emu
    cols = _case(ncols=2)
    del cols["msis"]
    p = dataclasses.replace(synth.StepParams(), use_msis=True)
    with pytest.raises(RuntimeError, match=_abi.ERRORS[4]):  # invalid arguments
        emu.advance(cols, p, solver=1)
