"""Host-side interface of the MI355X vertical step.

Mirrors GITM's ``advance_vertical_all`` (src/advance_vertical.f90:8-39): one
call advances every column it is given by one (or ``n_steps``) vertical
steps with the solver ``UseAUSMSolver`` selects (src/advance_vertical.f90
:157-161).  Stepping errors raise ``VerticalStepError``, the counterpart of the
reference's ``stop_gitm`` calls.

Arrays use the column-contiguous layout of :mod:`gitm_amd.synth`
(``[column][component][level]``, levels -1..nAlts+2).  ``advance`` works on
device-resident torch tensors (FP64, ``device="cuda"``, which is the HIP
device under ROCm) and queues the kernel on torch's current stream;
``advance_host`` takes numpy arrays and copies through the library.

The HIP library (``libgitm_vert.so``) must be built (``python -m
gitm_amd.build``); there is no fallback implementation.
"""
from __future__ import annotations

import ctypes
import dataclasses
import math
import os
import weakref

import numpy as np

from . import _abi
from .synth import StepParams

LIB_PATH = os.path.join(os.path.dirname(os.path.abspath(__file__)), "libgitm_vert.so")

STATE = ("nds", "vvel", "vel", "temp", "rho")
OUTPUTS = ("logns", "nden")
INPUTS = ("tempunit", "mmm", "gamma", "cmax", "eddy", "alt", "grav", "invr", "dalt",
          "coslat")
ION_FIELDS = ("ions", "ivel", "sza", "mlat")  # read/written with UseIonAdvection
# bottom boundary values: UseMSIS's level 0 per column; the tides' levels -1
# and 0 per column ([column][east, north, temp][level])
BOTTOM_FIELDS = ("msis", "tides")

_lib = None


class VerticalStepError(RuntimeError):
    def __init__(self, code: int, column: int = -1, level: int = 0):
        self.code, self.column, self.level = code, column, level
        where = f" (column {column}, level {level})" if column >= 0 else ""
        super().__init__(f"{_abi.ERRORS.get(code, code)}{where}")


def lib() -> ctypes.CDLL:
    """The loaded HIP library; raises if it has not been built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} is missing: build it with "
                               "`python -m gitm_amd.build`")
        # torch ships its own HIP runtime with the same soname
        # (libamdhip64.so.7).  Load it first so that the library binds to the
        # runtime torch's tensors and streams live in; one process must not
        # carry two HIP runtimes.
        try:
            import torch  # noqa: F401
        except ImportError:
            pass
        _lib = ctypes.CDLL(LIB_PATH)
        for name in ("gitm_vert_create", "gitm_vert_destroy", "gitm_vert_set_stream",
                     "gitm_vert_advance", "gitm_vert_advance_host", "gitm_vert_get_error",
                     "gitm_vert_set_tile"):
            getattr(_lib, name).restype = ctypes.c_int
        _lib.gitm_vert_strerror.restype = ctypes.c_char_p
        _lib.gitm_vert_prepare.restype = ctypes.c_int
        _lib.gitm_vert_prepare.argtypes = [ctypes.c_void_p, ctypes.POINTER(_abi.PrepParams),
                                           ctypes.POINTER(_abi.Columns),
                                           ctypes.POINTER(_abi.Field),
                                           ctypes.POINTER(_abi.Field), ctypes.c_void_p]
        vp, ci = ctypes.c_void_p, ctypes.c_int
        _lib.gitm_vert_batch_create.argtypes = [vp, ctypes.POINTER(_abi.Columns), ci, ci,
                                                ctypes.POINTER(vp)]
        _lib.gitm_vert_batch_destroy.argtypes = [vp]
        _lib.gitm_vert_batch_set_tile.argtypes = [vp, ctypes.POINTER(_abi.BatchTile)]
        _lib.gitm_vert_batch_get_tile.argtypes = [vp, ctypes.POINTER(_abi.BatchTile)]
        _lib.gitm_vert_batch_advance.argtypes = [vp, ctypes.POINTER(_abi.Params), ci, ci, vp]
        _lib.gitm_vert_batch_prepare.argtypes = [vp, ctypes.POINTER(_abi.PrepParams), vp]
        _lib.gitm_vert_batch_locate.argtypes = [vp, ci, ctypes.POINTER(ci),
                                                ctypes.POINTER(ci)]
        for name in ("gitm_vert_batch_create", "gitm_vert_batch_destroy",
                     "gitm_vert_batch_set_tile", "gitm_vert_batch_get_tile",
                     "gitm_vert_batch_advance",
                     "gitm_vert_batch_prepare", "gitm_vert_batch_locate"):
            getattr(_lib, name).restype = ctypes.c_int
        _lib.gitm_vert_sources_host.restype = ctypes.c_int
        _lib.gitm_vert_sources_host.argtypes = [vp, ctypes.POINTER(_abi.SourcesParams),
                                                ctypes.POINTER(_abi.Columns),
                                                ctypes.POINTER(_abi.SourcesFields)]
        _lib.gitm_vert_sources.restype = ctypes.c_int
        _lib.gitm_vert_sources.argtypes = [vp, ctypes.POINTER(_abi.SourcesParams),
                                           ctypes.POINTER(_abi.Columns),
                                           ctypes.POINTER(_abi.SourcesFields)]
        _lib.gitm_horiz_exchange.restype = ctypes.c_int
        _lib.gitm_horiz_exchange.argtypes = [vp, ctypes.POINTER(_abi.ExchangeParams),
                                             ctypes.POINTER(_abi.Columns),
                                             ctypes.POINTER(_abi.ExchangeFields)]
        _lib.gitm_horiz_advance_host.restype = ctypes.c_int
        _lib.gitm_horiz_advance_host.argtypes = [vp, ctypes.POINTER(_abi.HorizParams),
                                                 ctypes.POINTER(_abi.Columns),
                                                 ctypes.POINTER(_abi.HorizFields)]
        _lib.gitm_horiz_advance.restype = ctypes.c_int
        _lib.gitm_horiz_advance.argtypes = [vp, ctypes.POINTER(_abi.HorizParams),
                                            ctypes.POINTER(_abi.Columns),
                                            ctypes.POINTER(_abi.HorizFields)]
        _lib.gitm_vert_create.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_int]
        _lib.gitm_vert_destroy.argtypes = [ctypes.c_void_p]
        _lib.gitm_vert_set_stream.argtypes = [ctypes.c_void_p, ctypes.c_void_p]
        _lib.gitm_vert_set_tile.argtypes = [ctypes.c_void_p, ctypes.c_int]
        _lib.gitm_vert_get_error.argtypes = [ctypes.c_void_p, ctypes.POINTER(_abi.Error)]
        for name in ("gitm_vert_advance", "gitm_vert_advance_host"):
            getattr(_lib, name).argtypes = [ctypes.c_void_p, ctypes.POINTER(_abi.Params),
                                            ctypes.POINTER(_abi.Columns), ctypes.c_int,
                                            ctypes.c_int]
    return _lib


def make_params(sp: StepParams | None = None) -> _abi.Params:
    sp = sp or StepParams()
    p = _abi.Params()
    lib().gitm_vert_default_params(ctypes.byref(p))
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


def make_prep_params(pp=None) -> _abi.PrepParams:
    """C-ABI prep parameters from a :class:`gitm_amd.synth.PrepParams`."""
    from .synth import PrepParams
    pp = pp or PrepParams()
    p = _abi.PrepParams()
    lib().gitm_vert_default_prep_params(ctypes.byref(p))
    for s in range(_abi.NSPECIES):
        p.mass[s] = pp.mass[s]
        p.vibration[s] = pp.vibration[s]
    p.cfl, p.fixed_dt, p.dt_end = pp.cfl, pp.fixed_dt, pp.dt_end
    p.eddy_coef, p.eddy_p0, p.eddy_p1 = pp.eddy_coef, pp.eddy_p0, pp.eddy_p1
    p.use_ion_advection = int(pp.use_ion_advection)
    return p


def auto_waves(n_alts: int) -> int:
    """Wavefronts per column workgroup picked by default: the fewest 64-lane
    waves that hold nAlts+4 levels (one level per lane)."""
    return (n_alts + 4 + 63) // 64


def solver_id(solver) -> int:
    if isinstance(solver, str):
        return {"ausm": _abi.AUSM, "rusanov": _abi.RUSANOV}[solver.lower()]
    return int(solver)


def _device_columns(cols: dict, extra: tuple = ()):
    """C-ABI descriptor of device-resident columns (torch FP64 tensors in the
    column-contiguous layout) and the pointers of the ``extra`` fields."""
    import torch

    ptrs = {}
    for k, t in cols.items():
        if k not in _abi.COLUMN_KEYS and k not in extra:
            continue
        if t.dtype != torch.float64 or not t.is_cuda or not t.is_contiguous():
            raise ValueError(f"{k}: need a contiguous float64 device tensor")
        ptrs[k] = t.data_ptr()
    nc, L = cols["temp"].shape
    more = {k: ptrs.pop(k) for k in extra if k in ptrs}
    return _abi.columns_struct(ptrs, nc, L - 4), more


def _sources_structs(ptr: dict, L: int, dt, apply, use_turbulent_cond, mass):
    """SourcesFields and SourcesParams for column-contiguous arrays at the
    addresses ``ptr`` (keys as VerticalSolver.sources takes them)."""
    from .synth import species_masses

    names = {"raddist": "radial_distance", "vert_viscosity": "vertical_viscosity"}
    f = _abi.SourcesFields()
    for key, addr in ptr.items():
        name = names.get(key, key)
        ncomp = dict(_abi.SOURCES_FIELDS).get(name)
        if ncomp is not None:
            setattr(f, name, _abi.column_field(addr, ncomp, L))
    p = _abi.SourcesParams()
    p.dt = dt
    for s, m in enumerate(species_masses() if mass is None else mass):
        p.mass[s] = m
    p.use_turbulent_cond = int(use_turbulent_cond)
    p.apply = int(apply)
    return f, p


def _horiz_structs(ptr: dict, shape, dt, omega_body, beta_limiter, use_coriolis,
                   use_ion_advection, altitudes_per_workgroup, use_improved_ion_advection,
                   use_topography=False):
    """Columns, HorizFields and HorizParams for block-layout arrays at the
    addresses ``ptr`` (gitm_amd.synth.block_grid names)."""
    nb, L, NLg, NLo = (int(v) for v in shape)
    nlo, nla, na = NLo - 4, NLg - 4, L - 4

    def blk(key, ncomp=1):
        return _abi.gitm_block_field(ptr[key], nlo, nla, na, ncomp)

    c = _abi.Columns()
    c.n_lons, c.n_lats, c.n_blocks, c.n_alts = nlo, nla, nb, na
    c.rho, c.temperature, c.gamma = blk("rho"), blk("temp"), blk("gamma")
    c.velocity = blk("vel", 3)
    c.vertical_velocity = blk("vvel", _abi.NSPECIES)
    c.ndensity_s = blk("nds", _abi.NSPECIES)
    c.inv_radial_distance = blk("invr")
    c.cos_latitude = _abi.latitude_field(ptr["cos_lat"], nla)
    if use_ion_advection:
        c.idensity_s = blk("ions", _abi.NIONS)
        c.ivelocity = blk("ivel", 3)
    f = _abi.HorizFields()
    for key in ("inv_dlon_gb", "inv_dlon_fb", "dlon_fb", "inv_dlat_gb", "inv_dlat_fb",
                "dlat_fb"):
        setattr(f, key, blk(key))
    f.radial_distance = blk("raddist")
    f.latitude = _abi.latitude_field(ptr["latitude"], nla)
    f.tan_latitude = _abi.latitude_field(ptr["tan_lat"], nla)
    f.cmax_horizontal = _abi.gitm_block_field(ptr["cmax"], nlo, nla, na, 3, lon_lo=0,
                                              lat_lo=0, alt_lo=0, alt_hi=na + 1)
    if "temp_source" in ptr:
        f.temp_source = blk("temp_source")
    if use_ion_advection and use_improved_ion_advection:
        f.ivelocity_par = blk("ivel_par", 3)
        lo, la = ptr["grad_lon"], ptr["grad_lat"]
        f.grad_lon_m, f.grad_lon_0, f.grad_lon_p = (
            _abi.Field(lo + 8 * j * nlo, 1, 0, 3 * nlo, 0, 0, 0, 0) for j in range(3))
        f.grad_lat_m, f.grad_lat_0, f.grad_lat_p = (
            _abi.Field(la + 8 * j * nla, 0, 1, 3 * nla, 0, 0, 0, 0) for j in range(3))
    if use_topography:
        f.d_alt, f.dalt_dlon, f.dalt_dlat = blk("dalt"), blk("dalt_dlon"), blk("dalt_dlat")
    p = _abi.HorizParams()
    p.dt, p.omega_body, p.beta_limiter = dt, omega_body, beta_limiter
    p.use_topography = int(use_topography)
    p.use_coriolis, p.use_ion_advection = int(use_coriolis), int(use_ion_advection)
    p.use_improved_ion_advection = int(use_improved_ion_advection)
    p.altitudes_per_workgroup = altitudes_per_workgroup
    return c, f, p


class VerticalSolver:
    """A handle on one HIP device (``gitm_vert_create``)."""

    def __init__(self, device: int = -1, waves_per_column: int = 0):
        self._batches = weakref.WeakSet()  # destroyed before the handle
        self._h = ctypes.c_void_p()
        rc = lib().gitm_vert_create(ctypes.byref(self._h), device)
        if rc:
            raise VerticalStepError(rc)
        if waves_per_column:
            self.set_tile(waves_per_column)

    def close(self) -> None:
        for b in list(self._batches):
            b.close()
        if self._h:
            lib().gitm_vert_destroy(self._h)
            self._h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_tile(self, waves_per_column: int) -> None:
        rc = lib().gitm_vert_set_tile(self._h, waves_per_column)
        if rc:
            raise VerticalStepError(rc)

    def set_stream(self, stream_handle: int) -> None:
        lib().gitm_vert_set_stream(self._h, ctypes.c_void_p(stream_handle))

    def check(self) -> None:
        """Synchronise and raise the first stepping error, if any."""
        e = _abi.Error()
        rc = lib().gitm_vert_get_error(self._h, ctypes.byref(e))
        if rc:
            raise VerticalStepError(rc, e.column, e.level)

    # -- device-resident torch tensors -------------------------------------
    def advance(self, cols: dict, params: StepParams | None = None, solver="ausm",
                n_steps: int = 1, check: bool = True) -> None:
        """Advance device-resident columns in place (torch FP64 tensors).

        ``cols`` holds the STATE and INPUTS tensors (optionally "logns" and
        "nden" outputs) in the column-contiguous layout.
        """
        import torch

        c, _ = _device_columns(cols)
        p = make_params(params)
        self.set_stream(torch.cuda.current_stream().cuda_stream)
        rc = lib().gitm_vert_advance(self._h, ctypes.byref(p), ctypes.byref(c),
                                     solver_id(solver), n_steps)
        if rc:
            raise VerticalStepError(rc)
        if check:
            self.check()

    def prepare(self, cols: dict, pp=None, dt_out=None):
        """Column prep and the vertical CFL step on device tensors
        (``gitm_vert_prepare``): fills cols["nden"], "gamma", "mmm",
        "tempunit", "eddy", "cmax" (levels 0..nAlts+1) and, when present,
        "press" and "cp" from "nds", "temp", "rho", "vvel", "dalt" (and
        "ivel" with ion advection).  Returns a one-element device tensor with
        the time step (``dt_out`` if given); queued on the current stream."""
        import torch

        c, extra = _device_columns(cols, ("press", "cp"))
        L = cols["temp"].shape[1]
        pf = _abi.column_field(extra["press"], 1, L) if "press" in extra else None
        cpf = _abi.column_field(extra["cp"], 1, L) if "cp" in extra else None
        dt = dt_out if dt_out is not None else torch.empty(1, dtype=torch.float64,
                                                           device=cols["temp"].device)
        p = make_prep_params(pp)
        self.set_stream(torch.cuda.current_stream().cuda_stream)
        rc = lib().gitm_vert_prepare(self._h, ctypes.byref(p), ctypes.byref(c),
                                     ctypes.byref(pf) if pf is not None else None,
                                     ctypes.byref(cpf) if cpf is not None else None,
                                     ctypes.c_void_p(dt.data_ptr()))
        if rc:
            raise VerticalStepError(rc)
        return dt

    def sources(self, cols: dict, src: dict, dt: float, apply: bool = True,
                use_turbulent_cond: bool = False, mass=None) -> None:
        """The sources step's neutral terms (``gitm_vert_sources``) on device
        tensors: thermal conduction and viscosity from the state in ``cols``
        ("temp", "tempunit", "rho", "eddy", "nds", "vel", "vvel", "alt") and
        the coefficients in ``src`` ("kappa_temp", "visc_coef",
        "visc_coef_s", "cp", "raddist"; optional "heating", "chem_heating",
        "user_heating"); with ``apply`` the state is updated as add_sources
        does.  Increments land in src["conduction"], src["viscosity"],
        src["vert_viscosity"] when those tensors are present.  Queued on the
        current stream."""
        import torch

        c, _ = _device_columns(cols)
        for key, t in src.items():
            if t.dtype != torch.float64 or not t.is_cuda or not t.is_contiguous():
                raise ValueError(f"{key}: need a contiguous float64 device tensor")
        f, p = _sources_structs({k: t.data_ptr() for k, t in src.items()},
                                int(cols["temp"].shape[1]), dt, apply, use_turbulent_cond,
                                mass)
        self.set_stream(torch.cuda.current_stream().cuda_stream)
        rc = lib().gitm_vert_sources(self._h, ctypes.byref(p), ctypes.byref(c),
                                     ctypes.byref(f))
        if rc:
            raise VerticalStepError(rc)

    def sources_host(self, cols: dict, src: dict, dt: float, apply: bool = True,
                     use_turbulent_cond: bool = False, mass=None) -> None:
        """``sources`` on host numpy arrays (FP64, C-contiguous) through
        ``gitm_vert_sources_host``: staged on the device and back, in place,
        synchronous."""
        for k, a in list(cols.items()) + list(src.items()):
            if a.dtype != np.float64 or not a.flags.c_contiguous:
                raise ValueError(f"{k}: need a C-contiguous float64 array")
        ptrs = {k: a.ctypes.data for k, a in cols.items()
                if k in _abi.CANON or k in _abi.PER_COLUMN}
        nc, L = cols["temp"].shape
        c = _abi.columns_struct(ptrs, nc, L - 4)
        f, p = _sources_structs({k: a.ctypes.data for k, a in src.items()}, L, dt, apply,
                                use_turbulent_cond, mass)
        rc = lib().gitm_vert_sources_host(self._h, ctypes.byref(p), ctypes.byref(c),
                                          ctypes.byref(f))
        if rc:
            raise VerticalStepError(rc)

    def horizontal(self, grid: dict, dt: float, omega_body: float = 2.0 * math.pi / 86400.0,
                   beta_limiter: float = 2.0, use_coriolis: bool = True,
                   use_ion_advection: bool = False, altitudes_per_workgroup: int = 0,
                   use_improved_ion_advection: bool = False,
                   use_topography: bool = False) -> None:
        """The horizontal step (``gitm_horiz_advance``: advance_horizontal on
        every block) on device tensors in GITM's block layout
        ([block][component][level][lat][lon], halo cells included;
        gitm_amd.synth.block_grid names and shapes).  The state ("rho",
        "temp", "vel", "vvel", "nds", with ions "ions") is updated in place;
        "temp_source", when present, receives HorizontalTempSource.  Queued
        on the current stream; errors surface in ``check()``.  With ions and
        ``use_improved_ion_advection`` the divergence of "ivel_par" follows,
        with UAM_Gradient's coefficients "grad_lon" and "grad_lat"
        ([block][M, 0, P][cell]).  ``use_topography`` reads "dalt",
        "dalt_dlon" and "dalt_dlat" (state-shaped) and advances the levels
        in order, one launch each."""
        import torch

        for k, t in grid.items():
            if t.dtype != torch.float64 or not t.is_cuda or not t.is_contiguous():
                raise ValueError(f"{k}: need a contiguous float64 device tensor")
        c, f, p = _horiz_structs({k: t.data_ptr() for k, t in grid.items()},
                                 tuple(int(v) for v in grid["rho"].shape), dt, omega_body,
                                 beta_limiter, use_coriolis, use_ion_advection,
                                 altitudes_per_workgroup, use_improved_ion_advection,
                                 use_topography)
        self.set_stream(torch.cuda.current_stream().cuda_stream)
        rc = lib().gitm_horiz_advance(self._h, ctypes.byref(p), ctypes.byref(c),
                                      ctypes.byref(f))
        if rc:
            raise VerticalStepError(rc)

    def horizontal_host(self, grid: dict, dt: float, omega_body: float = 2.0 * math.pi / 86400.0,
                        beta_limiter: float = 2.0, use_coriolis: bool = True,
                        use_ion_advection: bool = False,
                        use_improved_ion_advection: bool = False,
                        use_topography: bool = False) -> None:
        """``horizontal`` on host numpy arrays (FP64, C-contiguous), through
        ``gitm_horiz_advance_host``: staged on the device and back, in place,
        synchronous; raises VerticalStepError on a failed step."""
        for k, a in grid.items():
            if a.dtype != np.float64 or not a.flags.c_contiguous:
                raise ValueError(f"{k}: need a C-contiguous float64 array")
        c, f, p = _horiz_structs({k: a.ctypes.data for k, a in grid.items()},
                                 grid["rho"].shape, dt, omega_body, beta_limiter,
                                 use_coriolis, use_ion_advection, 0,
                                 use_improved_ion_advection, use_topography)
        rc = lib().gitm_horiz_advance_host(self._h, ctypes.byref(p), ctypes.byref(c),
                                           ctypes.byref(f))
        if rc:
            raise VerticalStepError(rc)

    def exchange(self, grid: dict, n_blocks_lon: int, use_ion_advection: bool = False,
                 ion_velocity: str = "ivel_par") -> None:
        """The message exchange around the horizontal step
        (``gitm_horiz_exchange``) on device tensors in the block layout of
        ``horizontal``: the blocks form a regular grid ``n_blocks_lon``
        wide; "rho", "temp", "itemp", "etemp", "vel", "vvel", "nds" and,
        with ions, "ions" and the ion velocity ``grid[ion_velocity]`` get
        their halo cells; "latitude" marks the blocks at the poles.  Queued
        on the current stream; the checks surface in ``check()``."""
        import torch

        nb, L, NLg, NLo = (int(v) for v in grid["rho"].shape)
        nlo, nla, na = NLo - 4, NLg - 4, L - 4
        for k, t in grid.items():
            if t.dtype != torch.float64 or not t.is_cuda or not t.is_contiguous():
                raise ValueError(f"{k}: need a contiguous float64 device tensor")

        def blk(key, ncomp=1):
            return _abi.gitm_block_field(grid[key].data_ptr(), nlo, nla, na, ncomp)

        c = _abi.Columns()
        c.n_lons, c.n_lats, c.n_blocks, c.n_alts = nlo, nla, nb, na
        c.rho, c.temperature = blk("rho"), blk("temp")
        c.velocity = blk("vel", 3)
        c.vertical_velocity = blk("vvel", _abi.NSPECIES)
        c.ndensity_s = blk("nds", _abi.NSPECIES)
        f = _abi.ExchangeFields()
        f.i_temperature, f.e_temperature = blk("itemp"), blk("etemp")
        f.latitude = _abi.latitude_field(grid["latitude"].data_ptr(), nla)
        if ion_velocity in grid:
            f.ion_velocity = blk(ion_velocity, 3)
        if use_ion_advection:
            c.idensity_s = blk("ions", _abi.NIONS)
        p = _abi.ExchangeParams()
        p.n_blocks_lon, p.n_blocks_lat = n_blocks_lon, nb // max(n_blocks_lon, 1)
        p.use_ion_advection = int(use_ion_advection)
        self.set_stream(torch.cuda.current_stream().cuda_stream)
        rc = lib().gitm_horiz_exchange(self._h, ctypes.byref(p), ctypes.byref(c),
                                       ctypes.byref(f))
        if rc:
            raise VerticalStepError(rc)

    # -- host numpy arrays --------------------------------------------------
    def advance_host(self, cols: dict, params: StepParams | None = None,
                     solver="ausm", n_steps: int = 1) -> None:
        """Advance host columns in place (numpy FP64, C-contiguous)."""
        ptrs = {}
        for k, a in cols.items():
            if k not in _abi.COLUMN_KEYS:
                continue
            if a.dtype != np.float64 or not a.flags.c_contiguous:
                raise ValueError(f"{k}: need a C-contiguous float64 array")
            ptrs[k] = a.ctypes.data
        nc, L = cols["temp"].shape
        c = _abi.columns_struct(ptrs, nc, L - 4)
        p = make_params(params)
        rc = lib().gitm_vert_advance_host(self._h, ctypes.byref(p), ctypes.byref(c),
                                          solver_id(solver), n_steps)
        if rc:
            e = _abi.Error()
            raise VerticalStepError(rc)


class VerticalBatch:
    """Ensemble members with grids of their own, stepped together
    (``gitm_vert_batch_*``).

    ``members`` is a list of device column dicts (as for
    :meth:`VerticalSolver.advance`); their tensors must stay alive and in
    place for the life of the batch (the batch keeps references).  A step is
    one launch over every member's columns (its tile: :meth:`set_tile`), and
    ``advance``/``prepare`` allocate nothing, so a loop of them can be
    captured in a CUDA (HIP) graph.
    """

    TILE_KEYS = ("waves_per_workgroup", "columns_per_workgroup", "carries_in_lds",
                 "persistent")

    def __init__(self, solver: VerticalSolver, members: list, ions: bool = False):
        self.solver = solver
        self.members = list(members)
        n = len(self.members)
        descs = (_abi.Columns * n)()
        for i, cols in enumerate(self.members):
            descs[i], _ = _device_columns(cols)
        self.n_alts = sorted({int(m["temp"].shape[1]) - 4 for m in self.members})
        self._b = ctypes.c_void_p()
        rc = lib().gitm_vert_batch_create(solver._h, descs, n, int(ions),
                                          ctypes.byref(self._b))
        if rc:
            raise VerticalStepError(rc)
        solver._batches.add(self)

    def close(self) -> None:
        if self._b:
            lib().gitm_vert_batch_destroy(self._b)
            self._b = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_tile(self, tile=None) -> None:
        """Tile of the step (``gitm_vert_batch_tile``): a tuple or dict of
        (waves_per_workgroup, columns_per_workgroup, carries_in_lds,
        persistent); None for the automatic one."""
        if tile is None:
            rc = lib().gitm_vert_batch_set_tile(self._b, None)
        else:
            vals = [tile[k] for k in self.TILE_KEYS] if isinstance(tile, dict) else list(tile)
            rc = lib().gitm_vert_batch_set_tile(self._b, ctypes.byref(_abi.BatchTile(*vals)))
        if rc:
            raise VerticalStepError(rc)

    def tile(self) -> tuple:
        t = _abi.BatchTile()
        rc = lib().gitm_vert_batch_get_tile(self._b, ctypes.byref(t))
        if rc:
            raise VerticalStepError(rc)
        return tuple(getattr(t, k) for k in self.TILE_KEYS)

    def candidate_tiles(self) -> list:
        """Every tile the search tries: from the smallest workgroup that holds
        the tallest member up to four wavefronts, each count of short columns
        per workgroup, the RK4 carries where they fit, in LDS or in device
        memory, a grid of one workgroup per task or a persistent one (the
        library refuses the tiles the batch does not fit)."""
        w0 = auto_waves(max(self.n_alts))
        short = any(na + 4 <= 64 for na in self.n_alts)
        return [(w, c, carry, pers) for w in range(w0, 5)
                for c in (range(1, w + 1) if short else (w,))
                for carry in (2, 1, 0) for pers in (0, 1)]

    def locate(self, column: int) -> tuple:
        m, c = ctypes.c_int(), ctypes.c_int()
        rc = lib().gitm_vert_batch_locate(self._b, column, ctypes.byref(m), ctypes.byref(c))
        if rc:
            raise VerticalStepError(rc)
        return m.value, c.value

    def check(self) -> None:
        """Synchronise and raise the first stepping error, naming the member
        (``err.member``) and its column."""
        e = _abi.Error()
        rc = lib().gitm_vert_get_error(self.solver._h, ctypes.byref(e))
        if rc:
            member, col = self.locate(e.column) if e.column >= 0 else (-1, -1)
            err = VerticalStepError(rc, col, e.level)
            err.member = member
            raise err

    def advance(self, params: StepParams | None = None, solver="ausm", n_steps: int = 1,
                dt=None, check: bool = True) -> None:
        """Advance every member in place.  ``dt``: None (params.dt for all)
        or a device tensor of one step per member (e.g. from ``prepare``)."""
        import torch

        p = make_params(params)
        if dt is not None:
            if dt.dtype != torch.float64 or not dt.is_cuda or dt.numel() < len(self.members):
                raise ValueError("dt: need a float64 device tensor with one step per member")
        self.solver.set_stream(torch.cuda.current_stream().cuda_stream)
        rc = lib().gitm_vert_batch_advance(self._b, ctypes.byref(p), solver_id(solver),
                                           n_steps,
                                           ctypes.c_void_p(dt.data_ptr()) if dt is not None
                                           else None)
        if rc:
            raise VerticalStepError(rc)
        if check:
            self.check()

    def prepare(self, pp=None, dt_out=None):
        """Column prep of every member (``gitm_vert_batch_prepare``); returns
        a device tensor with each member's step (``dt_out`` if given)."""
        import torch

        n = len(self.members)
        dt = dt_out if dt_out is not None else torch.empty(
            n, dtype=torch.float64, device=self.members[0]["temp"].device)
        p = make_prep_params(pp)
        self.solver.set_stream(torch.cuda.current_stream().cuda_stream)
        rc = lib().gitm_vert_batch_prepare(self._b, ctypes.byref(p),
                                           ctypes.c_void_p(dt.data_ptr()))
        if rc:
            raise VerticalStepError(rc)
        return dt

    def autotune(self, params: StepParams | None = None, solver="ausm",
                 probe_columns: int = 16384, repeats: int = 3) -> dict:
        """Pick the step's tile by timing.

        A probe batch holds a copy of the leading columns of every member,
        about ``probe_columns`` in all and in the members' proportions, so it
        has the batch's mix of column heights.  It is stepped with every
        candidate tile (:meth:`candidate_tiles`) that the library accepts; the
        fastest is set on this batch.  The members' own state is not touched.
        Returns {tile: ms per probe step}.
        """
        import torch

        total = sum(int(m["temp"].shape[0]) for m in self.members)
        frac = min(1.0, probe_columns / total)
        probe = [{k: v[:max(1, int(v.shape[0] * frac))].clone() for k, v in m.items()}
                 for m in self.members]
        b = VerticalBatch(self.solver, probe,
                          ions=bool(params and params.use_ion_advection))
        times = {}
        try:
            for tile in self.candidate_tiles():
                try:
                    b.set_tile(tile)
                except VerticalStepError:
                    continue  # the members do not fit this tile
                b.advance(params, solver, 1)
                e0 = torch.cuda.Event(enable_timing=True)
                e1 = torch.cuda.Event(enable_timing=True)
                e0.record()
                for _ in range(repeats):
                    b.advance(params, solver, 1, check=False)
                e1.record()
                torch.cuda.synchronize()
                b.check()
                times[tile] = e0.elapsed_time(e1) / repeats
        finally:
            b.close()
        self.set_tile(min(times, key=times.get))
        return times


def to_device(cols: dict, device="cuda") -> dict:
    """Copy the state and input arrays of ``cols`` (and the ion fields when
    present) to the device, adding zeroed LogNS / NDensity outputs."""
    import torch

    out = {}
    for k in STATE + INPUTS + ION_FIELDS + BOTTOM_FIELDS:
        if k in cols:
            out[k] = torch.from_numpy(np.ascontiguousarray(cols[k],
                                                          dtype=np.float64)).to(device)
    out["logns"] = torch.zeros_like(out["nds"])
    out["nden"] = torch.zeros_like(out["temp"])
    return out


def to_host(cols: dict) -> dict:
    return {k: v.detach().cpu().numpy() for k, v in cols.items()}


def replace_params(sp: StepParams, **kw) -> StepParams:
    return dataclasses.replace(sp, **kw)
