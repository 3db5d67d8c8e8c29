"""The oracle's horizontal step (oh_advance_horizontal), pinned by the
reference build, its golden fixture and what the mathematics fixes.

* Against the reference Fortran advance_horizontal (compiled from its own
  sources by ``make -C oracle ref``, where they are present): bit for bit on
  seeded block grids with and without Coriolis, ion advection and the
  divergence of UseImprovedIonAdvection (UAM_Gradient, gradient.f90), and
  against the committed fixture written by that build.
* UAM_Gradient's stencils are exact for quadratics in each direction, the
  property its coefficients are built for (initialize.f90:268-292): the
  divergence of a field-aligned velocity quadratic in longitude, latitude
  and radius matches the analytic one.
* Linear density profiles under a uniform wind (zero temperature, so no
  pressure force changes the wind) are advected exactly: the averaged face
  values of linear data give the exact gradient, the Rusanov flux
  difference is 0, and RK4 of a constant rate is exact (cells far enough
  from the static halo cells that the four stages never see them).
* The Rusanov term alone lifts a V-shaped minimum (rise s per cell of
  width h on either side) at the closed-form rate cMax (1 - F1 + 2 F2) s / h:
  at the kink the MC limiter gives 0, next to it the mean of the 4th-order
  slopes, 2 (F1 - 2 F2) s / h, with the reference's F1 = 0.625 and
  F2 = 0.0416667.
* Uniform fields with a uniform east wind evolve only by the curvature,
  Coriolis and centrifugal terms and the tan(latitude) part of the
  divergence: the RK4 step matches a tight ODE integration of those terms.
* Halo cells and ghost levels are never written.
* UseTopography: bit for bit against the reference build; with level
  surfaces of zero slope it is the flat step bit for bit; and a profile
  linear in height, horizontally uniform, under a uniform east wind over
  sloping levels follows the closed-form recurrence in which each level
  reads the step's new value of the level below (the reference writes each
  level back before the next, advance_horizontal.f90:349-429), which
  updating every level from its input would miss by far.
"""
from __future__ import annotations

import os

import numpy as np
import pytest

from gitm_amd import synth
from oracle import cpu, ref

HERE = os.path.dirname(os.path.abspath(__file__))
OMEGA = 2.0 * np.pi / 86400.0
STATE = ("rho", "temp", "vel", "vvel", "nds")


@pytest.mark.skipif(not ref.horizontal_available(), reason="reference build absent")
@pytest.mark.parametrize("coriolis,ions,improved", [
    (True, True, False), (False, False, False), (True, False, False), (True, True, True),
    (False, True, True)])
def test_matches_reference_build(coriolis, ions, improved):
    g = synth.block_grid(synth.BlockGrid(2, 2, 50), seed=11 + int(ions), ions=ions)
    dt = synth.horizontal_dt(g, cfl=0.4)
    out, rc, _ = cpu.horizontal(g, dt, OMEGA, 2.0, coriolis, ions, improved)
    want = ref.horizontal(g, dt, OMEGA, 2.0, coriolis, ions, improved)
    assert rc == 0
    for k, v in want.items():
        got = out[k][:, 2:-2, 2:-2, 2:-2] if k == "temp_source" else out[k]
        assert np.array_equal(got, v), k


@pytest.mark.skipif(not ref.horizontal_available(), reason="reference build absent")
@pytest.mark.parametrize("ions", [False, True])
def test_topography_matches_reference_build(ions):
    g = synth.block_grid(synth.BlockGrid(2, 2, 50), seed=5 + int(ions), ions=ions,
                         topography=True)
    dt = synth.horizontal_dt(g, cfl=0.4)
    out, rc, _ = cpu.horizontal(g, dt, OMEGA, 2.0, True, ions, ions, use_topography=True)
    want = ref.horizontal(g, dt, OMEGA, 2.0, True, ions, ions, use_topography=True)
    flat, _, _ = cpu.horizontal(g, dt, OMEGA, 2.0, True, ions, ions)
    assert rc == 0
    assert not np.array_equal(out["nds"], flat["nds"])  # the term is not negligible
    for k, v in want.items():
        got = out[k][:, 2:-2, 2:-2, 2:-2] if k == "temp_source" else out[k]
        assert np.array_equal(got, v), k


def test_topography_with_level_surfaces_flat_is_flat_step():
    g = synth.block_grid(synth.BlockGrid(2, 1, 12, nlons=6, nlats=5), seed=3, ions=True,
                         topography=True)
    g["dalt_dlon"][:] = 0.0
    g["dalt_dlat"][:] = 0.0
    dt = synth.horizontal_dt(g, cfl=0.4)
    topo, rc1, _ = cpu.horizontal(g, dt, OMEGA, 2.0, True, True, True, use_topography=True)
    flat, rc2, _ = cpu.horizontal(g, dt, OMEGA, 2.0, True, True, True)
    assert rc1 == rc2 == 0
    for k in flat:
        assert np.array_equal(topo[k], flat[k]), k


def test_topography_levels_advance_in_order():
    nA = 10
    g = synth.block_grid(synth.BlockGrid(1, 1, nA), seed=2, topography=True)
    ve, slope, dalt, dt = 50.0, 1.0, 500.0, 10.0
    g["rho"][:] = 1.0
    g["temp"][:] = 0.0
    g["vel"][:] = 0.0
    g["vel"][:, 0] = ve
    g["nds"][:] = 1.0e15
    g["tan_lat"][:] = 0.0
    g["dalt"][:] = dalt
    g["dalt_dlon"][:] = slope
    g["dalt_dlat"][:] = -0.3
    lev = np.arange(nA + 4, dtype=np.float64)  # levels -1..nA+2
    for s in range(6):
        g["vvel"][:, s] = (10.0 + s + (3.0 - 0.4 * s) * lev)[None, :, None, None]
    out, rc, _ = cpu.horizontal(g, dt, OMEGA, 2.0, False, use_topography=True)
    assert rc == 0
    c = dt * ve * slope * (0.5 / dalt)
    for s in range(6):
        v = g["vvel"][0, s, :, 2, 2].copy()
        new = v.copy()
        for k in range(1, nA + 1):  # index k + 1
            new[k + 1] = v[k + 1] + c * (v[k + 2] - new[k])
        # cells 3 or more from the halo: the halo cells keep their input
        # values through the stages, so the cells next to them see a step,
        # which the level-to-level coupling carries inwards at about 2e-9;
        # reading the level below's input instead is off by about 0.1
        got = out["vvel"][0, s, 2:nA + 2, 4:-4, 4:-4]
        np.testing.assert_allclose(got, np.broadcast_to(new[2:nA + 2, None, None], got.shape),
                                   rtol=1e-7, atol=0)
        jacobi = v[2:nA + 2] + c * (v[3:nA + 3] - v[1:nA + 1])
        assert np.min(np.abs(new[3:nA + 2] - jacobi[1:])) > 0.1
    for k in ("rho", "vel", "nds"):  # nothing else moves
        assert np.array_equal(out[k], g[k]), k


def test_golden_fixture():
    import importlib.util
    spec = importlib.util.spec_from_file_location(
        "mkh", os.path.join(HERE, "golden", "make_golden_horizontal.py"))
    mk = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mk)
    fx = np.load(os.path.join(HERE, "golden", mk.NAME))
    g, dt = mk.inputs()
    assert str(fx["inputs_sha256"]) == mk.digest(g), "generator changed since the fixture"
    assert float(fx["dt"]) == dt
    out, rc, _ = cpu.horizontal(g, dt, mk.OMEGA, mk.BETA, True, True, True)
    assert rc == 0
    lv = fx["levels"]
    for k in STATE + ("ions",):
        assert np.array_equal(out[k][..., lv + 1, 2:-2, 2:-2], fx["out_" + k]), k
    assert np.array_equal(out["temp_source"][:, lv + 1, 2:-2, 2:-2], fx["out_temp_source"])


def uniform_grid(n=30, nalts=2, lat=0.0, h=1.0e5, r=6.7e6):
    """One n x n block of uniform geometry at one latitude."""
    L, N = nalts + 4, n + 4
    shp = (1, L, N, N)
    g = {"rho": np.full(shp, 1.0e-10), "temp": np.full(shp, 3.0e5),
         "vel": np.zeros((1, 3) + shp[1:]), "vvel": np.zeros((1, 6) + shp[1:]),
         "nds": np.full((1, 6) + shp[1:], 1.0e15), "gamma": np.full(shp, 1.5)}
    for k in ("inv_dlon_gb", "inv_dlon_fb", "inv_dlat_gb", "inv_dlat_fb"):
        g[k] = np.full(shp, 1.0 / h)
    g["dlon_fb"] = np.full(shp, h)
    g["dlat_fb"] = np.full(shp, h)
    g["invr"] = np.full(shp, 1.0 / r)
    g["raddist"] = np.full(shp, r)
    g["latitude"] = np.full((1, N), lat)
    g["tan_lat"] = np.tan(g["latitude"])
    g["cos_lat"] = np.cos(g["latitude"])
    g["cmax"] = np.full((1, 3, nalts + 2, n + 2, n + 2), 700.0)
    return g


def centre(a, n=30, m=3):
    """Cells at least 13 cells from the halos (four stages of 3-cell
    stencils never reach them): interior indices with halo offset 2."""
    c = slice(2 + 13, 2 + n - 13)
    return a[..., 2:-2, c, c] if m == 3 else a[..., c, c]


def test_linear_profiles_are_advected_exactly():
    n, h, U, V = 30, 1.0e5, 80.0, -50.0
    g = uniform_grid(n, h=h)
    x = (np.arange(n + 4) - 2.0) * h
    X, Y = np.meshgrid(x, x)  # [lat][lon]
    g["rho"][:] = 1.0e-10 * (1.0 + 2.0e-8 * X - 1.0e-8 * Y)
    for s in range(6):
        g["nds"][:, s] = 1.0e15 * (1.0 + 1.0e-8 * (s + 1) * X + 3.0e-9 * Y)
    g["vel"][:, 0] = U
    g["vel"][:, 1] = V
    g["temp"][:] = 0.0
    dt = 30.0
    out, rc, _ = cpu.horizontal(g, dt, OMEGA, 2.0, use_coriolis=False)
    assert rc == 0
    want = g["rho"] - dt * (U * 2.0e-8 * 1.0e-10 + V * -1.0e-8 * 1.0e-10)
    np.testing.assert_allclose(centre(out["rho"]), centre(want), rtol=2e-15, atol=0)
    for s in range(6):
        w = g["nds"][:, s] - dt * 1.0e15 * (U * 1.0e-8 * (s + 1) + V * 3.0e-9)
        np.testing.assert_allclose(centre(out["nds"][:, s]), centre(w), rtol=2e-15, atol=0)
    # no pressure force, no curvature at latitude 0: the winds stay uniform
    assert np.array_equal(centre(out["vel"]), centre(g["vel"]))
    assert np.array_equal(centre(out["temp"]), centre(g["temp"]))


def test_rusanov_term_lifts_a_minimum():
    n, h, c = 30, 1.0e5, 700.0
    g = uniform_grid(n, h=h)
    x = np.abs(np.arange(n + 4) - 2.0 - 14.0)  # V in longitude, minimum at cell 15
    s = 2.0e3  # K-units per cell
    g["temp"][:] = 3.0e5 + s * x[None, None, None, :]
    dt = 1.0e-3
    out, rc, _ = cpu.horizontal(g, dt, OMEGA, 2.0, use_coriolis=False)
    assert rc == 0
    rise = out["temp"][0, 2:-2, 2 + 14, 2 + 14] - g["temp"][0, 2:-2, 2 + 14, 2 + 14]
    F1, F2 = 0.625, 0.0416667
    np.testing.assert_allclose(rise, dt * c * (1.0 - F1 + 2.0 * F2) * s / h, rtol=1e-5)
    # a cell on the straight part barely moves (linear data: equal fluxes)
    far = out["temp"][0, 2:-2, 16, 2 + 20] - g["temp"][0, 2:-2, 16, 2 + 20]
    assert np.all(np.abs(far) < 1e-6 * np.abs(rise))


def test_source_terms_match_the_ode():
    from scipy.integrate import solve_ivp
    lat, r, U, gam = 0.6, 6.7e6, 120.0, 1.5
    g = uniform_grid(lat=lat, r=r)
    g["vel"][:, 0] = U
    g["gamma"][:] = gam
    dt = 200.0
    out, rc, _ = cpu.horizontal(g, dt, OMEGA, 2.0, use_coriolis=True)
    assert rc == 0
    tl, sl, cl = np.tan(lat), np.sin(lat), np.cos(lat)
    cs, cen = 2 * OMEGA * sl, OMEGA ** 2 * cl * sl

    def rhs(t, y):
        ve, vn, rho, T, nd = y
        div = -tl * vn / r
        return [-(ve * (-tl * vn) / r) + cs * vn,
                -(tl * ve * ve / r) - (cen * r + cs * ve),
                -rho * div, -(gam - 1) * T * div, -nd * div]

    y0 = [U, 0.0, 1.0e-10, 3.0e5, 1.0e15]
    sol = solve_ivp(rhs, (0, dt), y0, method="DOP853", rtol=1e-13, atol=1e-30)
    ve, vn, rho, T, nd = sol.y[:, -1]
    np.testing.assert_allclose(centre(out["vel"][:, 0]), ve, rtol=1e-10)
    np.testing.assert_allclose(centre(out["vel"][:, 1]), vn, rtol=1e-8)
    np.testing.assert_allclose(centre(out["rho"]), rho, rtol=1e-12)
    np.testing.assert_allclose(centre(out["temp"]), T, rtol=1e-12)
    np.testing.assert_allclose(centre(out["nds"][:, 0]), nd, rtol=1e-12)
    assert abs(vn) > 1.0  # the terms do act


def test_halos_and_ghost_levels_untouched():
    g = synth.block_grid(synth.BlockGrid(2, 1, 12, nlons=6, nlats=5), seed=3, ions=True)
    out, rc, _ = cpu.horizontal(g, synth.horizontal_dt(g), use_ion_advection=True)
    assert rc == 0
    for k in STATE + ("ions",):
        a, b = out[k], g[k]
        mask = np.ones(a.shape[-3:], bool)
        mask[2:-2, 2:-2, 2:-2] = False
        assert np.array_equal(a[..., mask], b[..., mask]), k
        assert not np.array_equal(a, b), k
    # electrons are the ion sum (advance_horizontal.f90:420-425)
    e = out["ions"][:, 9, 2:-2, 2:-2, 2:-2]
    acc = np.zeros_like(e)
    for s in range(9):
        acc = acc + out["ions"][:, s, 2:-2, 2:-2, 2:-2]
    assert np.array_equal(e, acc)


def test_negative_density_is_reported():
    g = synth.block_grid(synth.BlockGrid(1, 1, 10), seed=2)
    g["nds"][0, 3, 5, 6, 4] = -1.0e12  # level 4, lat 4, lon 2 (halo offset 2)
    _, rc, (col, lev) = cpu.horizontal(g, 1.0)
    assert rc == 1 and lev == 4 and col == (6 - 2) * 9 + (4 - 2)


def test_ion_divergence_of_quadratic_velocities():
    """With zero neutral and plain ion velocities the step leaves the ions
    alone except for the divergence term: n / (1 + Dt div), div from
    UAM_Gradient of a velocity quadratic in the cell indices, which its
    three-point stencils differentiate exactly."""
    g = synth.block_grid(synth.BlockGrid(2, 2, 20, nlons=6, nlats=5), seed=4, ions=True)
    for k in ("vel", "vvel", "ivel"):
        g[k][:] = 0.0
    g["cmax"][:] = 0.0  # no Rusanov term: the ions are uniform per level below
    g["ions"][:, :2] = 1.0e10
    nb, L, NLg, NLo = g["rho"].shape
    lon = np.arange(NLo) - 1.0
    lat = np.arange(NLg) - 1.0
    r = g["raddist"]
    LA, LO = np.meshgrid(lat, lon, indexing="ij")
    g["ivel_par"][:, 0] = (3.0 + 0.5 * LO + 0.02 * LO ** 2)[None, None]
    g["ivel_par"][:, 1] = (1.0 - 0.3 * LA + 0.04 * LA ** 2)[None, None]
    g["ivel_par"][:, 2] = 2.0 + 1e-5 * (r - 6.4e6) + 3e-12 * (r - 6.4e6) ** 2
    dt = 5.0
    out, rc, _ = cpu.horizontal(g, dt, use_coriolis=False, use_ion_advection=True,
                                use_improved_ion_advection=True)
    assert rc == 0
    # the analytic divergence in index space (the grid's longitudes and
    # latitudes are uniform, so d/dlon = (d/di) / dlon)
    lonr, latr = g["latitude"], g["cos_lat"]
    dlon = 2.0 * np.pi / (2 * 6)
    dlatr = np.pi / (2 * 5)
    i0 = np.s_[2:-2]
    for b in range(nb):
        for k in range(2, L - 2):
            rr = r[b, k, 2:-2, 2:-2]
            ir = 1.0 / rr
            LOi, LAi = LO[2:-2, 2:-2], LA[2:-2, 2:-2]
            cl = latr[b, 2:-2][:, None]
            tl = np.clip(g["tan_lat"][b, 2:-2][:, None], -5.0, 5.0)
            de = (0.5 + 0.04 * LOi) / dlon * ir / np.maximum(cl, 0.17)
            dn = (-0.3 + 0.08 * LAi) / dlatr * ir
            vn = 1.0 - 0.3 * LAi + 0.04 * LAi ** 2
            x = rr - 6.4e6
            du = 1e-5 + 6e-12 * x
            vu = 2.0 + 1e-5 * x + 3e-12 * x ** 2
            div = (de + dn - tl * vn * ir + du + 2 * vu * ir) * cl
            want = 1.0e10 / (1 + dt * div)
            np.testing.assert_allclose(out["ions"][b, 0, k, 2:-2, 2:-2], want, rtol=1e-12)
    # electrons re-summed on every cell, halo cells and ghost levels included
    e = np.zeros_like(out["ions"][:, 9])
    for s in range(9):
        e = e + out["ions"][:, s]
    assert np.array_equal(out["ions"][:, 9], e)
