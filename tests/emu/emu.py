"""TEST INFRASTRUCTURE ONLY: runs the host build of the column kernels."""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

from gitm_amd import _abi

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
LIB = os.path.join(HERE, "libgv_emu.so")
SRC = os.path.join(HERE, "emu_vert.cpp")
DEPS = [SRC] + [os.path.join(ROOT, "gitm_amd", "csrc", f)
                for f in ("vert_column.h", "vert_params.h", "vert_ad.h", "vert_adj.h")] + \
    [os.path.join(HERE, "count.h")] + \
    [os.path.join(ROOT, "include", "gitm_vert.h")]

_lib = None


def build(force: bool = False) -> None:
    if (not force and os.path.exists(LIB) and
            all(os.path.getmtime(LIB) >= os.path.getmtime(d) for d in DEPS)):
        return
    # built under a private name and renamed, so that concurrent test
    # processes never load a half-written library
    tmp = f"{LIB}.{os.getpid()}.tmp"
    subprocess.run(["g++", "-std=c++20", "-O2", "-fPIC", "-shared", "-ffp-contract=off",
                    "-pthread", "-I" + os.path.join(ROOT, "include"),
                    "-I" + os.path.join(ROOT, "gitm_amd", "csrc"), SRC, "-o", tmp],
                   check=True)
    os.replace(tmp, LIB)


def lib():
    global _lib
    if _lib is None:
        build()
        _lib = ctypes.CDLL(LIB)
        _lib.gv_emu_advance.restype = ctypes.c_int
        _lib.gv_emu_adjoint.restype = ctypes.c_int
        _lib.gv_emu_adjoint_hand.restype = ctypes.c_int
    return _lib


def params_from(sp) -> _abi.Params:
    p = _abi.Params()
    lib().gv_emu_default_params(ctypes.byref(p))
    p.dt = sp.dt
    p.omega_body = sp.omega_body
    p.max_vertical_velocity = sp.max_vert_vel
    p.beta_limiter = sp.beta_limiter
    p.min_tec = sp.min_tec
    p.vert_tau = sp.vert_tau
    p.use_neutral_friction = int(sp.use_neutral_friction)
    p.use_coriolis = int(sp.use_coriolis)
    p.use_ion_advection = int(sp.use_ion_advection)
    p.use_improved_ion_advection = int(sp.use_improved_ion_advection)
    p.use_nighttime_ion_bcs = int(sp.use_nighttime_ion_bcs)
    p.use_msis = int(getattr(sp, "use_msis", False))
    p.use_gswm_tides = int(getattr(sp, "use_gswm_tides", False))
    p.use_waccm_tides = int(getattr(sp, "use_waccm_tides", False))
    p.use_damping = int(getattr(sp, "use_damping", False))
    return p


def advance(cols: dict, sp, solver: int = 1, n_steps: int = 1, waves: int = 0,
            cols_range=None, device_math: bool = False) -> dict:
    """Run the host build of the column kernels.  ``device_math`` evaluates
    the expressions the GPU computes in a different form as the GPU does
    (friction powers as exp(e ln T), log-space sums); by default the
    reference's forms, so that the result can be compared with the oracle
    bit for bit."""
    nc, L = cols["temp"].shape
    na = L - 4
    out = {k: np.ascontiguousarray(np.array(cols[k], dtype=np.float64))
           for k in ("nds", "vvel", "vel", "temp", "rho")}
    out["logns"] = np.zeros_like(out["nds"])
    out["nden"] = np.zeros_like(out["temp"])
    keep = {k: np.ascontiguousarray(cols[k], dtype=np.float64)
            for k in ("tempunit", "mmm", "gamma", "cmax", "eddy", "alt", "grav", "invr",
                      "dalt", "coslat")}
    if sp.use_ion_advection:
        out["ions"] = np.ascontiguousarray(np.array(cols["ions"], dtype=np.float64))
        for k in ("ivel", "sza", "mlat"):
            keep[k] = np.ascontiguousarray(cols[k], dtype=np.float64)
    if getattr(sp, "use_msis", False) and "msis" in cols:
        keep["msis"] = np.ascontiguousarray(cols["msis"], dtype=np.float64)
    if "tides" in cols and (getattr(sp, "use_gswm_tides", False) or
                            getattr(sp, "use_waccm_tides", False)):
        keep["tides"] = np.ascontiguousarray(cols["tides"], dtype=np.float64)
    ptrs = {k: v.ctypes.data for k, v in out.items()}
    ptrs.update({k: v.ctypes.data for k, v in keep.items()})
    c = _abi.columns_struct(ptrs, nc, na)
    p = params_from(sp)
    e = _abi.Error()
    c0, c1 = cols_range if cols_range else (0, nc)
    rc = lib().gv_emu_advance(ctypes.byref(p), ctypes.byref(c), solver, n_steps, waves,
                              c0, c1, int(device_math), ctypes.byref(e))
    if rc:
        raise RuntimeError(f"emulated kernel: {_abi.ERRORS.get(rc, rc)} "
                           f"(column {e.column}, level {e.level})")
    return out


class AdjFields(ctypes.Structure):
    _fields_ = [(n, _abi.Field) for n in ("nds", "vvel", "vel", "temp", "rho")]


ADJ = (("nds", 6), ("vvel", 6), ("vel", 3), ("temp", 1), ("rho", 1))


def adj_fields(arrays: dict, L: int) -> AdjFields:
    a = AdjFields()
    for name, ncomp in ADJ:
        setattr(a, name, _abi.column_field(arrays[name].ctypes.data, ncomp, L))
    return a


def adjoint(cols: dict, sp, gout: dict, cols_range=None, nmax: int = 1 << 15,
            maxph: int = 1024):
    """Adjoint of one AUSM step (host build of the taped column code):
    ``gout`` holds the adjoints of the step's outputs (nds, vvel, vel, temp,
    rho); returns the adjoints of its inputs and the largest node and phase
    counts a column used."""
    nc, L = cols["temp"].shape
    keep = {k: np.ascontiguousarray(cols[k], dtype=np.float64)
            for k in ("nds", "vvel", "vel", "temp", "rho", "tempunit", "mmm", "gamma",
                      "cmax", "eddy", "alt", "grav", "invr", "dalt", "coslat")}
    c = _abi.columns_struct({k: v.ctypes.data for k, v in keep.items()}, nc, L - 4)
    go = {k: np.ascontiguousarray(gout[k], dtype=np.float64) for k, _ in ADJ}
    gi = {k: np.zeros_like(go[k]) for k, _ in ADJ}
    p = params_from(sp)
    e = _abi.Error()
    nodes, phases = ctypes.c_int(), ctypes.c_int()
    c0, c1 = cols_range if cols_range else (0, nc)
    rc = lib().gv_emu_adjoint(ctypes.byref(p), ctypes.byref(c),
                              ctypes.byref(adj_fields(go, L)), ctypes.byref(adj_fields(gi, L)),
                              c0, c1, nmax, maxph, ctypes.byref(nodes), ctypes.byref(phases),
                              ctypes.byref(e))
    if rc:
        raise RuntimeError(f"emulated adjoint: {_abi.ERRORS.get(rc, rc)} "
                           f"(column {e.column}, level {e.level})")
    return gi, nodes.value, phases.value


class AdjFixFields(ctypes.Structure):
    _fields_ = [(n, _abi.Field) for n in ("tempunit", "gamma", "mmm", "cmax", "eddy")]


def adjoint_hand(cols: dict, sp, gout: dict, cols_range=None, device_math: bool = False,
                 fixed: bool = False):
    """The hand-written adjoint of one AUSM step (host build of vert_adj.h):
    the adjoints of the step's inputs from those of its outputs (nds, vvel,
    vel, temp, rho); with ``fixed`` also those of the column-prep inputs
    (tempunit, gamma, mmm, cmax, eddy)."""
    nc, L = cols["temp"].shape
    keep = {k: np.ascontiguousarray(cols[k], dtype=np.float64)
            for k in ("nds", "vvel", "vel", "temp", "rho", "tempunit", "mmm", "gamma",
                      "cmax", "eddy", "alt", "grav", "invr", "dalt", "coslat")}
    c = _abi.columns_struct({k: v.ctypes.data for k, v in keep.items()}, nc, L - 4)
    go = {k: np.ascontiguousarray(gout[k], dtype=np.float64) for k, _ in ADJ}
    gi = {k: np.zeros_like(go[k]) for k, _ in ADJ}
    gf = {k: np.zeros((nc, L)) for k in ("tempunit", "gamma", "mmm", "cmax", "eddy")}
    fx = AdjFixFields()
    for k in gf:
        setattr(fx, k, _abi.column_field(gf[k].ctypes.data, 1, L))
    p = params_from(sp)
    e = _abi.Error()
    c0, c1 = cols_range if cols_range else (0, nc)
    rc = lib().gv_emu_adjoint_hand(ctypes.byref(p), ctypes.byref(c),
                                   ctypes.byref(adj_fields(go, L)), ctypes.byref(adj_fields(gi, L)),
                                   ctypes.byref(fx) if fixed else None, c0, c1,
                                   int(device_math), ctypes.byref(e))
    if rc:
        raise RuntimeError(f"emulated adjoint: {_abi.ERRORS.get(rc, rc)} "
                           f"(column {e.column}, level {e.level})")
    return (gi, gf) if fixed else gi


OP_CLASSES = ("add", "mul", "div", "sqrt", "trans", "cmp")


def count_ops(cols: dict, sp, solver: int = 1, col: int = 0, device_math: bool = True) -> dict:
    """FP64 operations of one step of column ``col`` (the column code over a
    counting number type): counts per class, summed over the lanes."""
    nc, L = cols["temp"].shape
    keep = {k: np.ascontiguousarray(cols[k], dtype=np.float64)
            for k in ("nds", "vvel", "vel", "temp", "rho", "tempunit", "mmm", "gamma",
                      "cmax", "eddy", "alt", "grav", "invr", "dalt", "coslat")}
    c = _abi.columns_struct({k: v.ctypes.data for k, v in keep.items()}, nc, L - 4)
    p = params_from(sp)
    out = (ctypes.c_longlong * 6)()
    rc = lib().gv_emu_count_ops(ctypes.byref(p), ctypes.byref(c), solver, col,
                                int(device_math), out)
    if rc:
        raise RuntimeError(f"operation count: {_abi.ERRORS.get(rc, rc)}")
    return dict(zip(OP_CLASSES, list(out)))
