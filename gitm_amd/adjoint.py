"""Reverse-mode sensitivities of the vertical step on the GPU.

``adjoint_step`` wraps ``gitm_vert_adjoint``: the adjoint of one AUSM step
(src/vertical_solver_ausm.f90:26-433 with the gather/scatter of
src/advance_vertical.f90:45-219), recomputed from the step's input state with
taped arithmetic and swept backwards on the device.  ``VerticalWindow`` is a
``torch.autograd.Function`` over a window of steps: the forward pass runs the
plain HIP step and keeps one checkpoint per step; the backward pass recomputes
each step from its checkpoint inside the adjoint kernel, last step first.

Between vertical steps the window applies a heating source the way
add_sources applies GITM's heating terms (src/add_sources.f90:53-66):
Temperature(1:nAlts) += Dt * Q / (cp * TempUnit), with Q in W/kg, cp from
calc_pressure and TempUnit from calc_rates (both held at their values at the
start of the window, as are the other column-prep inputs).

The HIP library must be built; there is no fallback.
"""
from __future__ import annotations

import ctypes

import numpy as np
import torch

from . import _abi, vertical
from .synth import StepParams

ADJ = (("nds", "ndensity_s", 6), ("vvel", "vertical_velocity", 6), ("vel", "velocity", 3),
       ("temp", "temperature", 1), ("rho", "rho", 1))
STATE = tuple(a for a, _, _ in ADJ)
FIXED = ("tempunit", "mmm", "gamma", "cmax", "eddy", "alt", "grav", "invr", "dalt", "coslat")


class AdjointFields(ctypes.Structure):
    _fields_ = [(c, _abi.Field) for _, c, _ in ADJ]


_bound = False


def _lib():
    global _bound
    lib = vertical.lib()
    if not _bound:
        lib.gitm_vert_adjoint.restype = ctypes.c_int
        lib.gitm_vert_adjoint.argtypes = [ctypes.c_void_p, ctypes.POINTER(_abi.Params),
                                          ctypes.POINTER(_abi.Columns),
                                          ctypes.POINTER(AdjointFields),
                                          ctypes.POINTER(AdjointFields)]
        lib.gitm_vert_set_adjoint_memory.restype = ctypes.c_int
        lib.gitm_vert_set_adjoint_memory.argtypes = [ctypes.c_void_p, ctypes.c_size_t]
        _bound = True
    return lib


def _fields(t: dict, L: int) -> AdjointFields:
    f = AdjointFields()
    for name, cname, ncomp in ADJ:
        x = t[name]
        if x.dtype != torch.float64 or not x.is_cuda or not x.is_contiguous():
            raise ValueError(f"{name}: need a contiguous float64 device tensor")
        setattr(f, cname, _abi.column_field(x.data_ptr(), ncomp, L))
    return f


def adjoint_step(solver: "vertical.VerticalSolver", cols: dict, gout: dict,
                 params: StepParams | None = None, gin: dict | None = None) -> dict:
    """Adjoints of one AUSM step's inputs from those of its outputs.

    ``cols``: the step's input state and column inputs (device tensors, not
    modified); ``gout``: d loss / d (nds, vvel, vel, temp, rho) after the
    step.  Returns d loss / d the same fields before the step (``gin`` if
    given; it may be ``gout``).
    """
    params = params or StepParams()
    if params.use_ion_advection:
        raise ValueError("the adjoint covers the neutral step (no ion advection)")
    c, _ = vertical._device_columns({k: cols[k] for k in STATE + FIXED})
    L = int(cols["temp"].shape[1])
    go = {k: gout[k].contiguous() for k in STATE}
    gi = gin if gin is not None else {k: torch.empty_like(go[k]) for k in STATE}
    p = vertical.make_params(params)
    solver.set_stream(torch.cuda.current_stream().cuda_stream)
    rc = _lib().gitm_vert_adjoint(solver._h, ctypes.byref(p), ctypes.byref(c),
                                  ctypes.byref(_fields(go, L)), ctypes.byref(_fields(gi, L)))
    if rc:
        raise vertical.VerticalStepError(rc)
    return gi


def set_adjoint_memory(solver: "vertical.VerticalSolver", nbytes: int) -> None:
    rc = _lib().gitm_vert_set_adjoint_memory(solver._h, nbytes)
    if rc:
        raise vertical.VerticalStepError(rc)


def heating_profile(alt, peak: float = 150.0e3, width: float = 20.0e3,
                    amplitude: float = 1.0e-5):
    """A heating rate in W/kg, Gaussian in altitude (levels as ``alt``)."""
    return amplitude * np.exp(-((np.asarray(alt) - peak) / width) ** 2)


def level_weights(alt: np.ndarray, z: float) -> np.ndarray:
    """Per-column linear-interpolation weights of the levels 1..nAlts that
    bracket altitude ``z`` (``alt``: [columns, levels -1..nAlts+2])."""
    nc, L = alt.shape
    w = np.zeros((nc, L))
    for c in range(nc):
        a = alt[c, 2:L - 2]
        i = int(np.searchsorted(a, z)) - 1
        i = min(max(i, 0), len(a) - 2)
        f = (z - a[i]) / (a[i + 1] - a[i])
        w[c, 2 + i] = 1.0 - f
        w[c, 3 + i] = f
    return w


def _source(temp, heat, cp_tu, dt: float):
    # Temperature(1:nAlts) += Dt * Q / (cp * TempUnit)   (add_sources.f90:53-66)
    out = temp.clone()
    out[:, 2:-2] = temp[:, 2:-2] + dt * heat[:, 2:-2] / cp_tu[:, 2:-2]
    return out


class VerticalWindow(torch.autograd.Function):
    """n_steps x (AUSM vertical step, heating source) with one checkpoint
    per step; differentiable in the initial state and the heating rate."""

    @staticmethod
    def forward(ctx, nds, vvel, vel, temp, rho, heat, solver, fixed, params, n_steps):
        params = params or StepParams()
        cols = {"nds": nds.detach().clone(), "vvel": vvel.detach().clone(),
                "vel": vel.detach().clone(), "temp": temp.detach().clone(),
                "rho": rho.detach().clone()}
        cols.update({k: fixed[k] for k in FIXED})
        cp_tu = fixed["cp"] * fixed["tempunit"]
        checkpoints = []
        for _ in range(n_steps):
            checkpoints.append({k: cols[k].clone() for k in STATE})
            solver.advance(cols, params, "ausm", 1)
            cols["temp"] = _source(cols["temp"], heat.detach(), cp_tu, params.dt)
        ctx.solver, ctx.fixed, ctx.params = solver, fixed, params
        ctx.checkpoints, ctx.cp_tu = checkpoints, cp_tu
        return tuple(cols[k] for k in STATE)

    @staticmethod
    def backward(ctx, *grads):
        lam = {}
        for k, g in zip(STATE, grads):
            ref = ctx.checkpoints[0][k]
            lam[k] = torch.zeros_like(ref) if g is None else g.contiguous().clone()
        dt = ctx.params.dt
        g_heat = torch.zeros_like(ctx.cp_tu)
        for ck in reversed(ctx.checkpoints):
            # the source step: d temp_out / d heat = Dt / (cp TempUnit)
            g_heat[:, 2:-2] += lam["temp"][:, 2:-2] * dt / ctx.cp_tu[:, 2:-2]
            cols = dict(ck)
            cols.update({k: ctx.fixed[k] for k in FIXED})
            lam = adjoint_step(ctx.solver, cols, lam, ctx.params, gin=lam)
        return (lam["nds"], lam["vvel"], lam["vel"], lam["temp"], lam["rho"], g_heat,
                None, None, None, None)


def window(state: dict, heat, solver, fixed: dict, params: StepParams | None = None,
           n_steps: int = 10) -> dict:
    """Differentiable window: the state (nds, vvel, vel, temp, rho) after
    ``n_steps`` x (AUSM step, heating source).  ``fixed`` holds the column
    inputs (FIXED) and "cp"."""
    out = VerticalWindow.apply(state["nds"], state["vvel"], state["vel"], state["temp"],
                               state["rho"], heat, solver, fixed, params, n_steps)
    return dict(zip(STATE, out))
