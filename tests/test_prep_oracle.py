"""The oracle's column prep and vertical time step (ov_prepare), pinned by
closed forms: calc_pressure's Gamma and cp (calc_pressure.f90:56-86),
calc_rates' mean mass (calc_rates.Earth.f90:34-50), the eddy-diffusion ramp
(Earth.f90:376-411) and calc_timestep_vertical (calc_timestep.f90:85-147)."""
from __future__ import annotations

import dataclasses

import numpy as np
import pytest

from gitm_amd import synth
from oracle import cpu

KB = 1.3807e-23


def _columns(nc=3, na=12, dens=None, temp=0.5, w=None, dalt=1000.0, rho=1e-9):
    L = na + 4
    nds = np.full((nc, 6, L), 1e14) if dens is None else np.broadcast_to(
        np.asarray(dens, dtype=float)[None, :, None], (nc, 6, L)).copy()
    vvel = np.zeros((nc, 6, L)) if w is None else np.asarray(w, dtype=float).copy()
    return dict(nds=nds, temp=np.full((nc, L), temp), rho=np.full((nc, L), rho),
                vvel=vvel, dalt=np.full((nc, L), dalt), ivel=np.zeros((nc, 3, L)))


@pytest.mark.parametrize("sp,vib", [(0, 5.0), (1, 7.0), (2, 7.0), (3, 5.0), (4, 7.0), (5, 5.0)])
def test_single_species_gamma_cp_mass(sp, vib):
    """A pure gas: Gamma = 1 + 2/(Vib-2) (5/3 atomic, 7/5 diatomic),
    cp = (Vib-2) k/(2m), MeanMajorMass = m n/(n+1), TempUnit = MMM/k."""
    dens = np.zeros(6)
    dens[sp] = 3e15
    out, _ = cpu.prepare(_columns(dens=dens), synth.PrepParams())
    m = synth.species_masses()[sp]
    assert np.allclose(out["gamma"], 1 + 2 / (vib - 2), rtol=1e-15, atol=0)
    assert np.allclose(out["cp"], (vib - 2) * KB / (2 * m), rtol=1e-15, atol=0)
    assert np.allclose(out["mmm"], m * 3e15 / (3e15 + 1), rtol=1e-15, atol=0)
    assert np.allclose(out["tempunit"], out["mmm"] / KB, rtol=1e-15, atol=0)
    assert np.all(out["nden"] == 3e15)


def test_empty_cell_mean_mass_falls_back_to_first_species():
    """Zero density gives MeanMajorMass = 0, replaced by Mass(1)
    (calc_rates.Earth.f90:42)."""
    out, _ = cpu.prepare(_columns(dens=np.zeros(6)), synth.PrepParams())
    assert np.all(out["mmm"] == synth.species_masses()[0])


def test_mixture_gamma_and_pressure():
    """Equal parts O and N2: Gamma = 1 + (1/3 + 1/5), Pressure = T rho."""
    dens = np.zeros(6)
    dens[0] = dens[2] = 1e15
    out, _ = cpu.prepare(_columns(dens=dens, temp=0.25, rho=4e-10), synth.PrepParams())
    assert np.allclose(out["gamma"], 1 + (1 / 3 + 1 / 5), rtol=1e-15, atol=0)
    assert np.all(out["press"] == 0.25 * 4e-10)


def test_eddy_ramp():
    """K = K0 above p0, linear between p1 and p0, zero below p1."""
    pp = synth.PrepParams()
    for P, want in ((0.02, 50.0), (0.0075, 25.0), (0.005, 0.0), (0.001, 0.0),
                    (0.01, 50.0 * (0.01 - 0.005) / (0.01 - 0.005))):
        out, _ = cpu.prepare(_columns(temp=1.0, rho=P), pp)
        assert np.allclose(out["eddy"], want, rtol=1e-15, atol=0), (P, out["eddy"][0, 0])


def test_cmax_and_cfl_step():
    """cMax = max_s |w_s| + sqrt(Gamma T) on levels 0..nAlts+1; the step is
    CFL * min over columns of dAlt/cMax, capped by FixedDt and DtEnd."""
    na, nc = 12, 3
    L = na + 4
    dens = np.zeros(6)
    dens[0] = 1e15                       # Gamma = 5/3
    w = np.zeros((nc, 6, L))
    w[1, 3, 7] = -900.0                  # column 1, level 6: |w| = 900
    w[2, 0, 2] = 50.0
    cols = _columns(nc=nc, na=na, dens=dens, temp=0.6, w=w, dalt=2000.0)
    pp = dataclasses.replace(synth.PrepParams(), fixed_dt=1e9)
    out, dt = cpu.prepare(cols, pp)
    cs = np.sqrt((5 / 3) * 0.6)
    assert np.allclose(out["cmax"][0, 1:L - 1], cs, rtol=1e-15, atol=0)
    assert out["cmax"][1, 7] == pytest.approx(900.0 + cs, rel=1e-15)
    assert out["cmax"][0, 0] == 0 and out["cmax"][0, L - 1] == 0  # untouched
    assert dt == pytest.approx(0.75 * 2000.0 / (900.0 + cs), rel=1e-15)
    # the caps
    assert cpu.prepare(cols, dataclasses.replace(pp, fixed_dt=0.5))[1] == 0.5
    assert cpu.prepare(cols, dataclasses.replace(pp, dt_end=0.25))[1] == 0.25
    # winds outside 1..nAlts enter cMax but not the step
    w2 = np.zeros((nc, 6, L))
    w2[:, 2, 1] = 5000.0                 # level 0
    out2, dt2 = cpu.prepare(dict(cols, vvel=w2), pp)
    assert out2["cmax"][0, 1] == pytest.approx(5000.0 + cs, rel=1e-15)
    assert dt2 == pytest.approx(0.75 * 2000.0 / cs, rel=1e-15)


def test_ion_cfl_limit():
    """With ion advection the ion speed |IVel(up)| + c_s limits the step too."""
    na, nc = 12, 2
    L = na + 4
    cols = _columns(nc=nc, na=na, temp=0.6, dalt=1500.0)
    cols["ivel"][1, 2, 5] = -3000.0
    pp = dataclasses.replace(synth.PrepParams(), fixed_dt=1e9)
    out, dt = cpu.prepare(cols, pp)
    g = out["gamma"][0, 5]
    _, dti = cpu.prepare(cols, dataclasses.replace(pp, use_ion_advection=True))
    assert dti == pytest.approx(0.75 * 1500.0 / (3000.0 + np.sqrt(g * 0.6)), rel=1e-15)
    assert dti < dt


def test_matches_the_generators_inputs():
    """The synthetic atmospheres' Gamma, MeanMajorMass, TempUnit and cMax are
    built with the same formulas (gitm_amd.synth); both restatements agree
    bit for bit."""
    cols = synth.make_columns(synth.CONFIGS["cpu_324x50"], seed=1)
    out, _ = cpu.prepare(cols, synth.PrepParams())
    for k in ("gamma", "mmm", "tempunit"):
        assert np.array_equal(out[k], cols[k]), k
    assert np.array_equal(out["cmax"][:, 1:-1], cols["cmax"][:, 1:-1])
