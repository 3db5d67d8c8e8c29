"""Reverse-mode sensitivities of the vertical step on the GPU.

``adjoint_step`` wraps ``gitm_vert_adjoint`` / ``gitm_vert_adjoint_all``: the
hand-written adjoint of one AUSM step (src/vertical_solver_ausm.f90:26-433
with the gather/scatter of src/advance_vertical.f90:45-219).  Each column
recomputes its step from the step's input state, recording every stage's
intermediates in device memory, and reverses the stages' phases on the device.
``taped=True`` takes the same adjoint through ``gitm_vert_adjoint_taped``
(every operation on a tape, swept back with atomics), the slow independent
check.  ``prepare_adjoint`` wraps ``gitm_vert_prepare_adjoint``, the adjoint
of the column prep, and ``source_adjoint`` ``gitm_vert_source_adjoint``, that
of the window's heating source.

``VerticalWindow`` is a ``torch.autograd.Function`` over a window of steps
that follows GITM's order within a time step: the column prep from the
state (calc_pressure, calc_rates' mean mass and TempUnit, the eddy
coefficient and cMax; ``gitm_vert_prepare``), the vertical step, then a
heating source applied the way add_sources applies GITM's heating terms
(src/add_sources.f90:53-66): Temperature(1:nAlts) += Dt * Q / (cp *
TempUnit), with Q in W/kg.  The forward pass keeps one checkpoint of the
state per step; the backward pass recomputes each step's prep from its
checkpoint, runs the step's adjoint with the prep inputs' adjoints, and
carries those back through the prep's adjoint, last step first.  The time
step is held at ``params.dt`` (GITM's FixedDt; the CFL limit of the prep is
not applied inside the window).  ``with_prep=False`` holds the prep outputs
at their values at the start of the window instead.

The HIP library must be built; there is no fallback.
"""
from __future__ import annotations

import ctypes

import numpy as np
import torch

from . import _abi, vertical
from .synth import PrepParams, StepParams

ADJ = (("nds", "ndensity_s", 6), ("vvel", "vertical_velocity", 6), ("vel", "velocity", 3),
       ("temp", "temperature", 1), ("rho", "rho", 1))
STATE = tuple(a for a, _, _ in ADJ)
PREP = ("tempunit", "gamma", "mmm", "cmax", "eddy")  # the step's prep inputs
GEOMETRY = ("alt", "grav", "invr", "dalt", "coslat")
FIXED = ("tempunit", "mmm", "gamma", "cmax", "eddy") + GEOMETRY


class AdjointFields(ctypes.Structure):
    _fields_ = [(c, _abi.Field) for _, c, _ in ADJ]


class AdjointPrepFields(ctypes.Structure):
    _fields_ = [(n, _abi.Field) for n in ("temp_unit", "gamma", "mean_major_mass", "cmax",
                                          "eddy")]


_bound = False


def _lib():
    global _bound
    lib = vertical.lib()
    if not _bound:
        vp = ctypes.c_void_p
        for name in ("gitm_vert_adjoint", "gitm_vert_adjoint_taped"):
            f = getattr(lib, name)
            f.restype = ctypes.c_int
            f.argtypes = [vp, ctypes.POINTER(_abi.Params), ctypes.POINTER(_abi.Columns),
                          ctypes.POINTER(AdjointFields), ctypes.POINTER(AdjointFields)]
        lib.gitm_vert_adjoint_all.restype = ctypes.c_int
        lib.gitm_vert_adjoint_all.argtypes = [vp, ctypes.POINTER(_abi.Params),
                                              ctypes.POINTER(_abi.Columns),
                                              ctypes.POINTER(AdjointFields),
                                              ctypes.POINTER(AdjointFields),
                                              ctypes.POINTER(AdjointPrepFields)]
        lib.gitm_vert_prepare_adjoint.restype = ctypes.c_int
        lib.gitm_vert_prepare_adjoint.argtypes = [vp, ctypes.POINTER(_abi.PrepParams),
                                                  ctypes.POINTER(_abi.Columns),
                                                  ctypes.POINTER(AdjointPrepFields),
                                                  ctypes.POINTER(_abi.Field),
                                                  ctypes.POINTER(AdjointFields)]
        dp = ctypes.POINTER(ctypes.c_double)
        lib.gitm_vert_source_adjoint.restype = ctypes.c_int
        lib.gitm_vert_source_adjoint.argtypes = [vp, ctypes.c_longlong, ctypes.c_int,
                                                 ctypes.c_double, dp, dp, dp, dp, dp, dp, dp]
        lib.gitm_vert_set_adjoint_memory.restype = ctypes.c_int
        lib.gitm_vert_set_adjoint_memory.argtypes = [vp, ctypes.c_size_t]
        lib.gitm_vert_release_adjoint_memory.restype = ctypes.c_int
        lib.gitm_vert_release_adjoint_memory.argtypes = [vp]
        _bound = True
    return lib


def _check(t, name):
    if t.dtype != torch.float64 or not t.is_cuda or not t.is_contiguous():
        raise ValueError(f"{name}: need a contiguous float64 device tensor")


def _fields(t: dict, L: int) -> AdjointFields:
    f = AdjointFields()
    for name, cname, ncomp in ADJ:
        _check(t[name], name)
        setattr(f, cname, _abi.column_field(t[name].data_ptr(), ncomp, L))
    return f


def _prep_fields(t: dict, L: int) -> AdjointPrepFields:
    f = AdjointPrepFields()
    for key, cname in zip(PREP, ("temp_unit", "gamma", "mean_major_mass", "cmax", "eddy")):
        _check(t[key], key)
        setattr(f, cname, _abi.column_field(t[key].data_ptr(), 1, L))
    return f


def adjoint_step(solver: "vertical.VerticalSolver", cols: dict, gout: dict,
                 params: StepParams | None = None, gin: dict | None = None,
                 taped: bool = False, gprep: dict | None = None) -> dict:
    """Adjoints of one AUSM step's inputs from those of its outputs.

    ``cols``: the step's input state and column inputs (device tensors, not
    modified); ``gout``: d loss / d (nds, vvel, vel, temp, rho) after the
    step.  Returns d loss / d the same fields before the step (``gin`` if
    given; it may be ``gout``).  With ``gprep`` (a dict of tensors shaped like
    "temp" under the keys of PREP) those receive d loss / d the step's prep
    inputs.  Queued on the current stream (the taped form synchronises).
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
    lib = _lib()
    args = [solver._h, ctypes.byref(p), ctypes.byref(c), ctypes.byref(_fields(go, L)),
            ctypes.byref(_fields(gi, L))]
    if taped:
        if gprep is not None:
            raise ValueError("the taped adjoint holds the prep inputs fixed")
        rc = lib.gitm_vert_adjoint_taped(*args)
    elif gprep is not None:
        rc = lib.gitm_vert_adjoint_all(*args, ctypes.byref(_prep_fields(gprep, L)))
    else:
        rc = lib.gitm_vert_adjoint(*args)
    if rc:
        raise vertical.VerticalStepError(rc)
    return gi


def prepare_adjoint(solver: "vertical.VerticalSolver", cols: dict, gprep: dict, gin: dict,
                    gcp=None, pp: PrepParams | None = None) -> dict:
    """Adds to ``gin`` (nds, vvel, temp, rho) the adjoints that the column
    prep's outputs (``gprep``: PREP keys; ``gcp``: cp, optional) carry back to
    the state in ``cols`` (the prep's input).  Queued on the current stream."""
    c, _ = vertical._device_columns({k: cols[k] for k in STATE + ("dalt", "gamma", "mmm",
                                                                  "tempunit", "eddy", "cmax")})
    L = int(cols["temp"].shape[1])
    p = vertical.make_prep_params(pp)
    cpf = None
    if gcp is not None:
        _check(gcp, "gcp")
        cpf = _abi.column_field(gcp.data_ptr(), 1, L)
    solver.set_stream(torch.cuda.current_stream().cuda_stream)
    rc = _lib().gitm_vert_prepare_adjoint(solver._h, ctypes.byref(p), ctypes.byref(c),
                                          ctypes.byref(_prep_fields(gprep, L)),
                                          ctypes.byref(cpf) if cpf is not None else None,
                                          ctypes.byref(_fields(gin, L)))
    if rc:
        raise vertical.VerticalStepError(rc)
    return gin


def source_adjoint(solver: "vertical.VerticalSolver", g_temp, cp, tempunit, heat, dt: float,
                   g_heat, g_tempunit=None, g_cp=None) -> None:
    """The adjoint of the heating source Temperature(1:nAlts) += Dt * Q /
    (cp * TempUnit) (``gitm_vert_source_adjoint``): from ``g_temp``, the
    temperature's adjoint after the source, ADDS a = g_temp Dt / (cp TempUnit)
    (levels 1..nAlts) to ``g_heat`` and, when given, -a Q / TempUnit to
    ``g_tempunit`` and writes -a Q / cp to ``g_cp``.  All [column][level]
    float64 device tensors.  Queued on the current stream."""
    if (g_tempunit is None) != (g_cp is None):
        raise ValueError("source_adjoint: g_tempunit and g_cp go together")
    ts = [g_temp, cp, tempunit, heat, g_heat] + ([g_tempunit, g_cp] if g_cp is not None else [])
    for i, t in enumerate(ts):
        _check(t, f"source_adjoint argument {i}")
        if t.shape != g_temp.shape:
            raise ValueError("source_adjoint: every tensor shaped like g_temp")
    ncol, L = int(g_temp.shape[0]), int(g_temp.shape[1])
    dp = ctypes.POINTER(ctypes.c_double)

    def ptr(t):
        return ctypes.cast(t.data_ptr(), dp) if t is not None else None

    solver.set_stream(torch.cuda.current_stream().cuda_stream)
    rc = _lib().gitm_vert_source_adjoint(solver._h, ncol, L - 4, float(dt), ptr(g_temp),
                                         ptr(cp), ptr(tempunit), ptr(heat), ptr(g_heat),
                                         ptr(g_tempunit), ptr(g_cp))
    if rc:
        raise vertical.VerticalStepError(rc)


def set_adjoint_memory(solver: "vertical.VerticalSolver", nbytes: int) -> None:
    rc = _lib().gitm_vert_set_adjoint_memory(solver._h, nbytes)
    if rc:
        raise vertical.VerticalStepError(rc)


def release_adjoint_memory(solver: "vertical.VerticalSolver") -> None:
    rc = _lib().gitm_vert_release_adjoint_memory(solver._h)
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


def _prep_buffers(like) -> dict:
    """Output tensors of the column prep, shaped like ``like``; zero, so that
    cMax's levels -1 and nAlts+2, which the prep does not write, stay zero
    when the buffers are reused."""
    return {k: torch.zeros_like(like) for k in PREP + ("cp", "nden")}


def _prep(solver, state: dict, geometry: dict, pp: PrepParams, bufs: dict | None = None) -> dict:
    """The column prep of ``state``: a dict with the prep outputs (PREP keys,
    "cp"), in ``bufs`` (from _prep_buffers; reused from step to step, they
    are overwritten) or in fresh tensors."""
    d = dict(state)
    d.update({k: geometry[k] for k in GEOMETRY})
    d.update(bufs if bufs is not None else _prep_buffers(state["temp"]))
    solver.prepare(d, pp)
    return {k: d[k] for k in PREP + ("cp",)}


class VerticalWindow(torch.autograd.Function):
    """n_steps x (column prep, AUSM vertical step, heating source) with one
    checkpoint per step; differentiable in the initial state and the heating
    rate.  ``fixed`` holds the geometry (GEOMETRY keys) and, with
    ``with_prep=False``, the prep outputs (PREP keys and "cp") to hold."""

    @staticmethod
    def forward(ctx, nds, vvel, vel, temp, rho, heat, solver, fixed, params, n_steps,
                with_prep, prep_params):
        params = params or StepParams()
        pp = prep_params or PrepParams()
        state = {"nds": nds.detach().clone(), "vvel": vvel.detach().clone(),
                 "vel": vel.detach().clone(), "temp": temp.detach().clone(),
                 "rho": rho.detach().clone()}
        geometry = {k: fixed[k] for k in GEOMETRY}
        held = None if with_prep else {k: fixed[k] for k in PREP + ("cp",)}
        checkpoints = []
        bufs = _prep_buffers(state["temp"]) if with_prep else None
        for _ in range(n_steps):
            checkpoints.append({k: state[k].clone() for k in STATE})
            prep = _prep(solver, state, geometry, pp, bufs) if with_prep else held
            cols = dict(state)
            cols.update(geometry)
            cols.update({k: prep[k] for k in PREP})
            solver.advance(cols, params, "ausm", 1, check=False)
            state = {k: cols[k] for k in STATE}
            cp_tu = prep["cp"] * prep["tempunit"]
            state["temp"] = _source(state["temp"], heat.detach(), cp_tu, params.dt)
        solver.check()
        ctx.solver, ctx.geometry, ctx.held, ctx.params, ctx.pp = solver, geometry, held, params, pp
        ctx.checkpoints, ctx.heat = checkpoints, heat.detach().contiguous()
        return tuple(state[k] for k in STATE)

    @staticmethod
    def backward(ctx, *grads):
        lam = {}
        for k, g in zip(STATE, grads):
            ref = ctx.checkpoints[0][k]
            lam[k] = torch.zeros_like(ref) if g is None else g.contiguous().clone()
        # the adjoints before the step (the step's adjoint reads lam, the
        # adjoints after the step and the source, and writes nxt)
        nxt = {k: torch.empty_like(lam[k]) for k in STATE}
        dt = ctx.params.dt
        heat = ctx.heat
        g_heat = torch.zeros_like(heat)
        gprep = {k: torch.empty_like(heat) for k in PREP}
        gcp = torch.empty_like(heat) if ctx.held is None else None
        bufs = _prep_buffers(heat) if ctx.held is None else None
        for ck in reversed(ctx.checkpoints):
            prep = _prep(ctx.solver, ck, ctx.geometry, ctx.pp, bufs) if ctx.held is None \
                else ctx.held
            cols = dict(ck)
            cols.update(ctx.geometry)
            cols.update({k: prep[k] for k in PREP})
            # the source (temp += Dt Q / (cp TempUnit) on 1..nAlts) passes
            # the temperature's adjoint through unchanged; its cp and
            # TempUnit come from the same prep as the step's
            if ctx.held is None:
                adjoint_step(ctx.solver, cols, lam, ctx.params, gin=nxt, gprep=gprep)
                source_adjoint(ctx.solver, lam["temp"], prep["cp"], prep["tempunit"], heat, dt,
                               g_heat, gprep["tempunit"], gcp)
                prepare_adjoint(ctx.solver, cols, gprep, nxt, gcp=gcp, pp=ctx.pp)
            else:
                adjoint_step(ctx.solver, cols, lam, ctx.params, gin=nxt)
                source_adjoint(ctx.solver, lam["temp"], prep["cp"], prep["tempunit"], heat, dt,
                               g_heat)
            lam, nxt = nxt, lam
        ctx.solver.check()
        return (lam["nds"], lam["vvel"], lam["vel"], lam["temp"], lam["rho"], g_heat,
                None, None, None, None, None, None)


def window(state: dict, heat, solver, fixed: dict, params: StepParams | None = None,
           n_steps: int = 10, with_prep: bool = True,
           prep_params: PrepParams | None = None) -> dict:
    """Differentiable window: the state (nds, vvel, vel, temp, rho) after
    ``n_steps`` x (column prep, AUSM step, heating source).  ``fixed`` holds
    the geometry (GEOMETRY) and, with ``with_prep=False``, the prep outputs
    to hold (PREP and "cp")."""
    out = VerticalWindow.apply(state["nds"], state["vvel"], state["vel"], state["temp"],
                               state["rho"], heat, solver, fixed, params, n_steps, with_prep,
                               prep_params)
    return dict(zip(STATE, out))
