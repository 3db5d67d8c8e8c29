"""The vertical step on GITM's own block layout, as the Fortran drop-in
describes it (module arrays (-1:nLons+2,-1:nLats+2,-1:nAlts+2[,nComp],nBlocks)
in column-major order, NDensityS with all nSpeciesTotal components,
cMax_GDB(0:nLons+1,0:nLats+1,0:nAlts+1,3,nBlocks), KappaEddyDiffusion
without halo cells, CosLatitude(-1:nLats+2,nBlocks), per-block copies of
TempUnit and MeanMajorMass): the same columns as the column-contiguous
layout, advanced through the descriptors, give the same bits; the halo cells,
filled with NaN, are never read or written."""
from __future__ import annotations

import ctypes

import numpy as np
import pytest

from gitm_amd import _abi, synth

pytestmark = pytest.mark.gpu

NLON = NLAT = 9
NSP_TOTAL = 11  # nSpeciesTotal: the step uses the first six


def to_blocks(cols, nb, L):
    """Column arrays [col][comp][level] -> GITM block arrays, C order
    [block][comp][level][lat][lon] (== Fortran (lon, lat, level, comp, block))."""
    NLo, NLg = NLON + 4, NLAT + 4
    c = np.arange(nb * NLON * NLAT)
    i, j, b = c % NLON, (c // NLON) % NLAT, c // (NLON * NLAT)

    def halo(field, ncomp_alloc=None):
        a = np.asarray(cols[field])
        a3 = a[:, None, :] if a.ndim == 2 else a
        nc = a3.shape[1] if ncomp_alloc is None else ncomp_alloc
        out = np.full((nb, nc, L, NLg, NLo), np.nan)
        for comp in range(a3.shape[1]):
            out[b, comp, :, j + 2, i + 2] = a3[:, comp, :]
        return out

    blk = {"nds": halo("nds", NSP_TOTAL), "vvel": halo("vvel"), "vel": halo("vel"),
           "temp": halo("temp"), "rho": halo("rho"), "gamma": halo("gamma"),
           "alt": halo("alt"), "grav": halo("grav"), "invr": halo("invr"),
           "dalt": halo("dalt"), "tempunit": halo("tempunit"), "mmm": halo("mmm")}
    blk["logns"] = np.full((nb, 6, L, NLg, NLo), np.nan)
    blk["nden"] = np.full((nb, 1, L, NLg, NLo), np.nan)
    cm = np.full((nb, 3, L - 2, NLAT + 2, NLON + 2), np.nan)  # (0:n+1, 0:m+1, 0:nA+1, 3)
    cm[b, 2, :, j + 1, i + 1] = np.asarray(cols["cmax"])[:, 1:L - 1]
    blk["cmax"] = cm
    ed = np.full((nb, L, NLAT, NLON), np.nan)  # (nLons, nLats, -1:nAlts+2)
    ed[b, :, j, i] = np.asarray(cols["eddy"])
    blk["eddy"] = ed
    cl = np.full((nb, NLg), np.nan)  # (-1:nLats+2, nBlocks)
    cl[b, j + 2] = np.asarray(cols["coslat"])
    blk["coslat"] = cl
    return {k: np.ascontiguousarray(v) for k, v in blk.items()}, (i, j, b)


def from_blocks(a, idx, ncomp):
    i, j, b = idx
    return np.stack([a[b, comp, :, j + 2, i + 2] for comp in range(ncomp)], axis=1)


@pytest.mark.parametrize("which", ["ausm", "rusanov"])
def test_block_layout_matches_column_layout(solver, which):
    import torch
    from gitm_amd import vertical

    nb, nalts = 4, 30
    L = nalts + 4
    cols = synth.make_columns(synth.GridSpec(2, 2, nalts), seed=13)
    p = synth.StepParams()
    ref = vertical.to_device({k: np.array(v) for k, v in cols.items()}, "cuda")
    solver.advance(ref, p, which, 2)
    ref = vertical.to_host(ref)

    blk, idx = to_blocks(cols, nb, L)
    d = {k: torch.from_numpy(v).cuda() for k, v in blk.items()}
    c = _abi.Columns()
    c.n_lons, c.n_lats, c.n_blocks, c.n_alts = NLON, NLAT, nb, nalts

    def f(key, nc):
        return _abi.gitm_block_field(d[key].data_ptr(), NLON, NLAT, nalts, nc)

    c.ndensity_s = f("nds", NSP_TOTAL)
    c.vertical_velocity = f("vvel", 6)
    c.velocity = f("vel", 3)
    c.log_ns = f("logns", 6)
    for key, name in (("temp", "temperature"), ("rho", "rho"), ("nden", "ndensity"),
                      ("tempunit", "temp_unit"), ("mmm", "mean_major_mass"),
                      ("gamma", "gamma"), ("alt", "altitude"), ("grav", "gravity"),
                      ("invr", "inv_radial_distance"), ("dalt", "dalt")):
        setattr(c, name, f(key, 1))
    cm = _abi.gitm_block_field(d["cmax"].data_ptr(), NLON, NLAT, nalts, 3, comp0=2,
                               lon_lo=0, lat_lo=0, alt_lo=0, alt_hi=nalts + 1)
    c.cmax = cm
    c.eddy = _abi.gitm_block_field(d["eddy"].data_ptr(), NLON, NLAT, nalts, 1, lon_lo=1,
                                   lat_lo=1)
    c.cos_latitude = _abi.latitude_field(d["coslat"].data_ptr(), NLAT)
    params = vertical.make_params(p)
    solver.set_stream(torch.cuda.current_stream().cuda_stream)
    rc = vertical.lib().gitm_vert_advance(solver._h, ctypes.byref(params), ctypes.byref(c),
                                          vertical.solver_id(which), 2)
    assert rc == 0
    torch.cuda.synchronize()
    solver.check()
    out = {k: v.cpu().numpy() for k, v in d.items()}
    for key, nc in (("nds", 6), ("vvel", 6), ("vel", 3), ("temp", 1), ("rho", 1),
                    ("logns", 6), ("nden", 1)):
        got = from_blocks(out[key], idx, nc)
        want = ref[key] if ref[key].ndim == 3 else ref[key][:, None, :]
        assert np.array_equal(got, want), key
    # halo cells and the species beyond the sixth: never touched
    hm = np.ones((NLAT + 4, NLON + 4), bool)
    hm[2:-2, 2:-2] = False
    for key in ("nds", "vvel", "vel", "temp", "rho", "logns", "nden"):
        assert np.all(np.isnan(out[key][..., hm])), key
    assert np.all(np.isnan(out["nds"][:, 6:]))


def test_host_entry_on_block_layout(solver):
    """gitm_vert_advance_host on the same host arrays (what the Fortran shim
    calls): the bits of the device entry on the column layout."""
    from gitm_amd import vertical

    nb, nalts = 4, 30
    L = nalts + 4
    cols = synth.make_columns(synth.GridSpec(2, 2, nalts), seed=14)
    p = synth.StepParams()
    ref = vertical.to_device({k: np.array(v) for k, v in cols.items()}, "cuda")
    solver.advance(ref, p, "ausm", 1)
    ref = vertical.to_host(ref)
    blk, idx = to_blocks(cols, nb, L)
    c = _abi.Columns()
    c.n_lons, c.n_lats, c.n_blocks, c.n_alts = NLON, NLAT, nb, nalts

    def f(key, nc, **kw):
        return _abi.gitm_block_field(blk[key].ctypes.data, NLON, NLAT, nalts, nc, **kw)

    c.ndensity_s = f("nds", NSP_TOTAL)
    c.vertical_velocity, c.velocity, c.log_ns = f("vvel", 6), f("vel", 3), f("logns", 6)
    for key, name in (("temp", "temperature"), ("rho", "rho"), ("nden", "ndensity"),
                      ("tempunit", "temp_unit"), ("mmm", "mean_major_mass"),
                      ("gamma", "gamma"), ("alt", "altitude"), ("grav", "gravity"),
                      ("invr", "inv_radial_distance"), ("dalt", "dalt")):
        setattr(c, name, f(key, 1))
    c.cmax = f("cmax", 3, comp0=2, lon_lo=0, lat_lo=0, alt_lo=0, alt_hi=nalts + 1)
    c.eddy = f("eddy", 1, lon_lo=1, lat_lo=1)
    c.cos_latitude = _abi.latitude_field(blk["coslat"].ctypes.data, NLAT)
    params = vertical.make_params(p)
    rc = vertical.lib().gitm_vert_advance_host(solver._h, ctypes.byref(params),
                                               ctypes.byref(c), 1, 1)
    assert rc == 0
    for key, nc in (("nds", 6), ("vvel", 6), ("vel", 3), ("temp", 1), ("rho", 1),
                    ("logns", 6), ("nden", 1)):
        want = ref[key] if ref[key].ndim == 3 else ref[key][:, None, :]
        assert np.array_equal(from_blocks(blk[key], idx, nc), want), key
