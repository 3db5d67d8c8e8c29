"""The oracle's message exchange (oh_exchange), pinned by geometry.

Every halo cell of a block sits at a point of its block's extended
longitude/latitude grid; past a pole that point is the sphere's point at
the mirrored latitude and the longitude turned by 180 degrees.  So:
* scalar fields that are functions of the Cartesian position on the sphere,
  evaluated at the extended coordinates, are what the exchange must put in
  every halo cell (faces and corners) on levels 1..nAlts;
* the east and north components of a rigid rotation (W x r), evaluated in
  the extended frame, are what it must put there for the winds: past a pole
  the extended east and north unit vectors are the negatives of the true
  ones, which is the sign change of exchange_messages_sphere.f90:30-72;
* on the ghost levels, which the exchange does not carry, only that sign
  change happens, in the pole rows;
* Temperature < 100 and Rho < 0 on the blocks' cells are reported.
The reference exchange itself (ModSphereInterface.f90 over the
srcSphereAB/ block tree and MPI) cannot run here; these properties pin the
restatement instead."""
from __future__ import annotations

import numpy as np
import pytest

from gitm_amd import synth
from oracle import cpu


def sphere_grid(nbl, nbt, nalts=4, nlo=6, nla=5, ions=True):
    spec = synth.BlockGrid(nbl, nbt, nalts, nlons=nlo, nlats=nla)
    g = synth.block_grid(spec, seed=1, ions=ions)
    nb = spec.nblocks
    dlon = 2.0 * np.pi / (nbl * nlo)
    bi = np.arange(nb) % nbl
    lon = (bi[:, None] * nlo + np.arange(-1, nlo + 3)[None, :] - 0.5) * dlon  # [nb][NLo]
    lat = g["latitude"]                                                       # [nb][NLg]
    LON = lon[:, None, :]
    LAT = lat[:, :, None]
    x, y, z = np.cos(LAT) * np.cos(LON), np.cos(LAT) * np.sin(LON), np.sin(LAT) + 0 * LON
    L = nalts + 4
    lev = np.arange(L)[None, :, None, None]
    xs, ys, zs = x[:, None], y[:, None], z[:, None]

    def scal(a, b, c, d):
        return (a + b * xs + c * ys * zs + d * zs ** 2) * (1.0 + 0.01 * lev)

    want = {"rho": scal(2.0, 0.3, 0.2, 0.1), "temp": scal(5e4, 3e3, 1e3, 2e3),
            "itemp": scal(6e4, 1e3, 2e3, 3e3), "etemp": scal(7e4, 2e3, 3e3, 1e3)}
    want["nds"] = np.stack([scal(1.0 + s, 0.1 * s, 0.2, 0.05) for s in range(6)], axis=1)
    want["vvel"] = np.stack([scal(0.1 * s, 1.0, -0.5, 0.3) for s in range(6)], axis=1)
    want["ions"] = np.stack([scal(3.0 + s, 0.2, 0.1 * s, 0.4) for s in range(10)], axis=1)

    def rotation(W):
        # east and north components of W x r in the (extended) local frame
        ex = np.stack([-np.sin(LON), np.cos(LON), 0 * LON], -1)
        nx = np.stack([-np.sin(LAT) * np.cos(LON), -np.sin(LAT) * np.sin(LON),
                       np.cos(LAT) + 0 * LON], -1)
        r = np.stack([x, y, z], -1)
        V = np.cross(np.broadcast_to(W, r.shape), r)
        u, v = (V * ex).sum(-1), (V * nx).sum(-1)
        up = 0.1 * zs[:, 0] + 0 * u
        return np.stack([u, v, up], axis=1)[:, :, None] * (1.0 + 0.01 * lev)

    want["vel"] = rotation(np.array([30.0, -20.0, 80.0]))
    want["ivel"] = rotation(np.array([-50.0, 10.0, 40.0]))
    want = {k: np.ascontiguousarray(np.broadcast_to(v, v.shape)) for k, v in want.items()}
    grid = {k: v.copy() for k, v in want.items()}
    grid["latitude"] = lat
    return grid, want


def halo_mask(NLg, NLo):
    m = np.ones((NLg, NLo), bool)
    m[2:-2, 2:-2] = False
    return m


@pytest.mark.parametrize("nbl,nbt,ions", [(4, 3, True), (1, 2, False), (6, 1, True)])
def test_halos_are_the_sphere_fields(nbl, nbt, ions):
    grid, want = sphere_grid(nbl, nbt, ions=ions)
    nb, L, NLg, NLo = grid["rho"].shape
    hm = halo_mask(NLg, NLo)
    inp = {k: v.copy() for k, v in grid.items()}
    for k in cpu.EXCHANGED + ("ions", "ivel"):
        inp[k][..., 2:L - 2, :, :][..., hm] = 1.0e30  # the halos of levels 1..nAlts
    out, rc, _ = cpu.exchange(inp, nbl, ions)
    assert rc == 0
    keys = cpu.EXCHANGED + (("ions", "ivel") if ions else ())
    for k in keys:
        got, w = out[k][..., 2:L - 2, :, :], want[k][..., 2:L - 2, :, :]
        np.testing.assert_allclose(got[..., hm], w[..., hm], rtol=1e-12, atol=1e-9, err_msg=k)
        # the blocks' own cells are untouched
        assert np.array_equal(out[k][..., 2:-2, 2:-2], inp[k][..., 2:-2, 2:-2]), k


def test_ghost_levels_only_change_sign_in_pole_rows():
    grid, _ = sphere_grid(4, 2)
    nb, L, NLg, NLo = grid["rho"].shape
    out, rc, _ = cpu.exchange(grid, 4, True)
    assert rc == 0
    for k in ("vel", "ivel"):
        for lev in (0, 1, L - 2, L - 1):  # levels -1, 0, nAlts+1, nAlts+2
            a, b = out[k][:, :, lev], grid[k][:, :, lev]
            for blk in range(nb):
                south, north = blk < 4, blk >= 4
                rows = [0, 1] if south else [NLg - 2, NLg - 1]
                other = [r for r in range(NLg) if r not in rows]
                np.testing.assert_array_equal(a[blk, :2][:, rows], -b[blk, :2][:, rows])
                np.testing.assert_array_equal(a[blk, 2][:, rows], b[blk, 2][:, rows])
                np.testing.assert_array_equal(a[blk][:, other], b[blk][:, other])
        assert not np.array_equal(out[k], grid[k])
    for k in ("rho", "temp", "nds"):  # scalars: ghost levels untouched
        for lev in (0, 1, L - 2, L - 1):
            np.testing.assert_array_equal(out[k][..., lev, :, :], grid[k][..., lev, :, :])


def test_checks_after_the_exchange():
    grid, _ = sphere_grid(2, 2, ions=False)
    g = {k: v.copy() for k, v in grid.items()}
    g["temp"][3, 2, 4, 5] = 50.0  # block 3, level 1, lat 3, lon 4
    _, rc, (col, lev) = cpu.exchange(g, 2)
    assert rc == 5 and lev == 1 and col == (3 * 5 + 2) * 6 + 3
    g = {k: v.copy() for k, v in grid.items()}
    g["rho"][1, 4, 2, 2] = -1.0
    _, rc, (col, lev) = cpu.exchange(g, 2)
    assert rc == 1 and lev == 3 and col == (1 * 5 + 0) * 6 + 0
