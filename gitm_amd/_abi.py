"""ctypes mirror of include/gitm_vert.h.

Describes arrays to the C ABI.  ``column_major_fields`` builds descriptors for
the package's column-contiguous layout ([column][component][level], level
fastest); ``gitm_block_field`` builds one for a GITM module array
(-1:nLons+2, -1:nLats+2, -1:nAlts+2[, nComp], nBlocks) in Fortran order.
"""
from __future__ import annotations

import ctypes

NSPECIES = 6
NIONS = 10

RUSANOV = 0
AUSM = 1

ERRORS = {
    0: "ok",
    1: "non-positive neutral density entering a column",
    2: "singular neutral-friction matrix",
    3: "negative ion density too close to the upper boundary",
    4: "invalid arguments",
    5: "unsupported option (BC perturbation, UseDamping with the AUSM solver)",
    6: "HIP runtime error",
    7: "out of device memory",
    8: "adjoint tape too small for the column",
    9: "temperature below 100 after the message exchange",
    10: "columns too large for the fixed scratch of gitm_vert_advance_noalloc",
}


class Field(ctypes.Structure):
    _fields_ = [("ptr", ctypes.c_void_p),
                ("s_lon", ctypes.c_int64), ("s_lat", ctypes.c_int64),
                ("s_blk", ctypes.c_int64), ("s_lev", ctypes.c_int64),
                ("s_comp", ctypes.c_int64),
                ("first_level", ctypes.c_int32), ("reserved", ctypes.c_int32)]


FIELD_NAMES = ("ndensity_s", "vertical_velocity", "velocity", "temperature", "rho",
               "log_ns", "ndensity", "temp_unit", "mean_major_mass", "gamma",
               "cmax", "eddy", "altitude", "gravity", "inv_radial_distance", "dalt",
               "cos_latitude", "idensity_s", "ivelocity", "sza", "mlat", "msis_bc",
               "tides_east", "tides_north", "tides_temp")


class Columns(ctypes.Structure):
    _fields_ = [("n_lons", ctypes.c_int32), ("n_lats", ctypes.c_int32),
                ("n_blocks", ctypes.c_int32), ("n_alts", ctypes.c_int32)] + \
        [(n, Field) for n in FIELD_NAMES]


class Params(ctypes.Structure):
    _fields_ = [("dt", ctypes.c_double), ("omega_body", ctypes.c_double),
                ("max_vertical_velocity", ctypes.c_double),
                ("beta_limiter", ctypes.c_double), ("min_tec", ctypes.c_double),
                ("vert_tau", ctypes.c_double),
                ("mass", ctypes.c_double * NSPECIES),
                ("thermal_diff_coef", ctypes.c_double * NSPECIES),
                ("use_neutral_friction", ctypes.c_int32),
                ("use_coriolis", ctypes.c_int32),
                ("use_ion_advection", ctypes.c_int32),
                ("use_improved_ion_advection", ctypes.c_int32),
                ("use_nighttime_ion_bcs", ctypes.c_int32),
                ("use_msis", ctypes.c_int32), ("use_gswm_tides", ctypes.c_int32),
                ("use_waccm_tides", ctypes.c_int32),
                ("use_bc_perturbation", ctypes.c_int32),
                ("during_perturb", ctypes.c_int32), ("use_damping", ctypes.c_int32),
                ("reserved", ctypes.c_int32)]


class PrepParams(ctypes.Structure):
    _fields_ = [("mass", ctypes.c_double * NSPECIES),
                ("vibration", ctypes.c_double * NSPECIES),
                ("cfl", ctypes.c_double), ("fixed_dt", ctypes.c_double),
                ("dt_end", ctypes.c_double), ("eddy_coef", ctypes.c_double),
                ("eddy_p0", ctypes.c_double), ("eddy_p1", ctypes.c_double),
                ("use_ion_advection", ctypes.c_int32), ("reserved", ctypes.c_int32)]


class Error(ctypes.Structure):
    _fields_ = [("code", ctypes.c_int32), ("column", ctypes.c_int32),
                ("level", ctypes.c_int32), ("reserved", ctypes.c_int32)]


class BatchTile(ctypes.Structure):
    _fields_ = [("waves_per_workgroup", ctypes.c_int32),
                ("columns_per_workgroup", ctypes.c_int32),
                ("carries_in_lds", ctypes.c_int32), ("persistent", ctypes.c_int32)]


# canonical field name -> (C-ABI field, components)
CANON = {
    "nds": ("ndensity_s", NSPECIES), "vvel": ("vertical_velocity", NSPECIES),
    "vel": ("velocity", 3), "temp": ("temperature", 1), "rho": ("rho", 1),
    "logns": ("log_ns", NSPECIES), "nden": ("ndensity", 1),
    "tempunit": ("temp_unit", 1), "mmm": ("mean_major_mass", 1),
    "gamma": ("gamma", 1), "cmax": ("cmax", 1), "eddy": ("eddy", 1),
    "alt": ("altitude", 1), "grav": ("gravity", 1),
    "invr": ("inv_radial_distance", 1), "dalt": ("dalt", 1),
    "ions": ("idensity_s", NIONS), "ivel": ("ivelocity", 3),
}
PER_COLUMN = {"coslat": "cos_latitude", "sza": "sza", "mlat": "mlat"}
# per-column fields with components ([column][component])
PER_COLUMN_MULTI = {"msis": ("msis_bc", 9)}
# the bottom tides, [column][east, north, temp][level -1, 0]
TIDES = ("tides_east", "tides_north", "tides_temp")
# every key columns_struct takes
COLUMN_KEYS = set(CANON) | set(PER_COLUMN) | set(PER_COLUMN_MULTI) | {"tides"}


def column_field(ptr: int, ncomp: int, nlev: int) -> Field:
    """Descriptor of a [column][component][level] array (levels -1..nAlts+2)."""
    return Field(ptr, ncomp * nlev, 0, 0, 1, nlev, -1, 0)


def per_column_field(ptr: int) -> Field:
    return Field(ptr, 1, 0, 0, 0, 0, -1, 0)


def columns_struct(ptrs: dict, n_cols: int, n_alts: int) -> Columns:
    """Columns struct for arrays in the column-contiguous layout.

    ``ptrs`` maps canonical names (see CANON / PER_COLUMN) to data addresses.
    """
    c = Columns()
    c.n_lons, c.n_lats, c.n_blocks, c.n_alts = n_cols, 1, 1, n_alts
    L = n_alts + 4
    for name, addr in ptrs.items():
        if name in CANON:
            fname, ncomp = CANON[name]
            setattr(c, fname, column_field(addr, ncomp, L))
        elif name in PER_COLUMN:
            setattr(c, PER_COLUMN[name], per_column_field(addr))
        elif name in PER_COLUMN_MULTI:
            fname, ncomp = PER_COLUMN_MULTI[name]
            setattr(c, fname, Field(addr, ncomp, 0, 0, 0, 1, 0, 0))
        elif name == "tides":
            for i, fname in enumerate(TIDES):
                setattr(c, fname, Field(addr + 16 * i, 6, 0, 0, 1, 0, -1, 0))
        else:
            raise KeyError(name)
    return c


def gitm_block_field(ptr: int, n_lons: int, n_lats: int, n_alts: int, ncomp_alloc: int,
                     comp0: int = 0, lon_lo: int = -1, lat_lo: int = -1,
                     alt_lo: int = -1, alt_hi: int | None = None, blocks: bool = True,
                     lon_n: int | None = None, lat_n: int | None = None) -> Field:
    """Descriptor of a Fortran array A(lon_lo:.., lat_lo:.., alt_lo:alt_hi[, comp], blk)
    whose first element is at ``ptr``; the descriptor points at (1,1,alt_lo,comp0+1,1).
    """
    if alt_hi is None:
        alt_hi = n_alts + 2
    lon_n = lon_n if lon_n is not None else (n_lons + 4 if lon_lo == -1 else
                                             n_lons + 2 if lon_lo == 0 else n_lons)
    lat_n = lat_n if lat_n is not None else (n_lats + 4 if lat_lo == -1 else
                                             n_lats + 2 if lat_lo == 0 else n_lats)
    s_lon = 1
    s_lat = lon_n
    s_lev = lon_n * lat_n
    nlev = alt_hi - alt_lo + 1
    s_comp = s_lev * nlev
    s_blk = s_comp * max(ncomp_alloc, 1) if blocks else 0
    first = ptr + 8 * ((1 - lon_lo) * s_lon + (1 - lat_lo) * s_lat + comp0 * s_comp)
    return Field(first, s_lon, s_lat, s_blk, s_lev, s_comp, alt_lo, 0)


class SourcesParams(ctypes.Structure):
    _fields_ = [("dt", ctypes.c_double), ("mass", ctypes.c_double * NSPECIES),
                ("use_turbulent_cond", ctypes.c_int32), ("apply", ctypes.c_int32)]


SOURCES_FIELDS = (("kappa_temp", 1), ("visc_coef", 1), ("visc_coef_s", NSPECIES), ("cp", 1),
                  ("radial_distance", 1), ("heating", 1), ("chem_heating", 1),
                  ("user_heating", 1), ("conduction", 1), ("viscosity", 3),
                  ("vertical_viscosity", NSPECIES))


class SourcesFields(ctypes.Structure):
    _fields_ = [(n, Field) for n, _ in SOURCES_FIELDS]


class HorizParams(ctypes.Structure):
    _fields_ = [("dt", ctypes.c_double), ("omega_body", ctypes.c_double),
                ("beta_limiter", ctypes.c_double), ("use_coriolis", ctypes.c_int32),
                ("use_ion_advection", ctypes.c_int32),
                ("use_improved_ion_advection", ctypes.c_int32),
                ("use_topography", ctypes.c_int32),
                ("altitudes_per_workgroup", ctypes.c_int32), ("reserved", ctypes.c_int32)]


HORIZ_FIELDS = ("inv_dlon_gb", "inv_dlon_fb", "dlon_fb", "inv_dlat_gb", "inv_dlat_fb",
                "dlat_fb", "radial_distance", "latitude", "tan_latitude", "cmax_horizontal",
                "temp_source", "ivelocity_par", "grad_lon_m", "grad_lon_0", "grad_lon_p",
                "grad_lat_m", "grad_lat_0", "grad_lat_p", "d_alt", "dalt_dlon", "dalt_dlat")


class HorizFields(ctypes.Structure):
    _fields_ = [(n, Field) for n in HORIZ_FIELDS]


def latitude_field(ptr: int, n_lats: int) -> Field:
    """Descriptor of a Latitude-like array (-1:nLats+2, nBlocks): ptr at (1,1)."""
    return Field(ptr + 8 * 2, 0, 1, n_lats + 4, 0, 0, 0, 0)


class ExchangeParams(ctypes.Structure):
    _fields_ = [("n_blocks_lon", ctypes.c_int32), ("n_blocks_lat", ctypes.c_int32),
                ("use_ion_advection", ctypes.c_int32), ("reserved", ctypes.c_int32)]


class ExchangeFields(ctypes.Structure):
    _fields_ = [(n, Field) for n in ("i_temperature", "e_temperature", "ion_velocity",
                                     "latitude")]
