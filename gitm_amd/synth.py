"""Seeded synthetic atmospheres for the vertical step.

Builds the column working set that GITM's ``advance_vertical`` reads
(src/advance_vertical.f90:84-151) for a grid of independent columns, in the
column-contiguous layout used throughout this package: every per-level field
is stored ``[column][component][level]`` with the level index fastest and
levels -1..nAlts+2 at offsets 0..nAlts+3.

The column prep mirrors the reference's own formulas:

* uniform altitudes ``AltMin + (iAlt-0.5)*dz``  (src/initialize.f90:138-141);
* ``dAlt``, ``1/r`` and ``g = -g0*(R/r)**2``     (src/initialize.f90:148-170);
* ``Gamma`` and ``cp``                           (src/calc_pressure.f90:56-86);
* ``MeanMajorMass = sum(m n)/(N+1)``, ``TempUnit = MMM/k``
                                                 (src/calc_rates.Earth.f90:33-50);
* ``cMax = max|w_s| + sqrt(Gamma*T)``            (src/calc_timestep.f90:114-118);
* eddy coefficient pressure ramp                 (src/Earth.f90:376-411).

Random numbers come from a counter-based generator (splitmix64 + Box-Muller),
so a (seed, stream, index) triple always yields the same value.
"""
from __future__ import annotations

import dataclasses
import math

import numpy as np

NSPECIES = 6
NIONS = 10
KB = 1.3807e-23          # ModConstants.f90:15
AMU = 1.6726e-27         # ModConstants.f90:11
R_EARTH = 6372.0 * 1000.0  # ModEarth.f90:73
VIBRATION = (5.0, 7.0, 7.0, 5.0, 7.0, 5.0)  # ModEarth.f90:307-312
BASE_DENSITY = (1e17, 1e18, 4e18, 1e13, 1e14, 1e14)  # O O2 N2 N NO He at 100 km

STATE_FIELDS = ("nds", "vvel", "vel", "temp", "rho")
PREP_FIELDS = ("tempunit", "mmm", "gamma", "cmax", "eddy", "alt", "grav",
               "invr", "dalt")
COMPONENTS = {"nds": NSPECIES, "vvel": NSPECIES, "vel": 3, "logns": NSPECIES,
              "ions": NIONS, "ivel": 3}


@dataclasses.dataclass
class PrepParams:
    """Inputs of the column prep and the vertical time step: the reference
    defaults (ModInputs.f90:123,302-304; ModEarth.f90:271-281,307-312)."""
    mass: tuple = dataclasses.field(default_factory=lambda: tuple(species_masses()))
    vibration: tuple = VIBRATION
    cfl: float = 0.75
    fixed_dt: float = 2.0
    dt_end: float = 1.0e32
    eddy_coef: float = 50.0
    eddy_p0: float = 0.01
    eddy_p1: float = 0.005
    use_ion_advection: bool = False


def species_masses() -> np.ndarray:
    """Species masses in kg, formed as init_planet does (ModEarth.f90:271-281)."""
    mo = 16.0 * AMU
    mn = 14.0 * AMU
    return np.array([mo, 2.0 * mo, 2.0 * mn, mn, mn + mo, 4.0 * AMU])


# ---------------------------------------------------------------------------
# counter-based random numbers
# ---------------------------------------------------------------------------
_M64 = np.uint64(0xFFFFFFFFFFFFFFFF)


def _splitmix64(x: np.ndarray) -> np.ndarray:
    with np.errstate(over="ignore"):
        z = x + np.uint64(0x9E3779B97F4A7C15)
        z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
        z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
        return z ^ (z >> np.uint64(31))


def _uniform(seed: int, stream: int, n: int) -> np.ndarray:
    """n doubles in (0, 1) from counter (seed, stream, 0..n-1)."""
    base = _splitmix64(np.array([(seed * 1000003 + stream) & 0xFFFFFFFFFFFFFFFF],
                                dtype=np.uint64))[0]
    with np.errstate(over="ignore"):
        ctr = np.arange(n, dtype=np.uint64) + base
    bits = _splitmix64(ctr) >> np.uint64(11)
    return (bits.astype(np.float64) + 0.5) * (1.0 / 9007199254740992.0)


def normal(seed: int, stream: int, shape) -> np.ndarray:
    """Standard normals (Box-Muller) from counter (seed, stream)."""
    n = int(np.prod(shape))
    m = (n + 1) // 2
    u1 = _uniform(seed, 2 * stream, m)
    u2 = _uniform(seed, 2 * stream + 1, m)
    r = np.sqrt(-2.0 * np.log(u1))
    z = np.concatenate([r * np.cos(2 * np.pi * u2), r * np.sin(2 * np.pi * u2)])
    return z[:n].reshape(shape)


# ---------------------------------------------------------------------------
@dataclasses.dataclass
class GridSpec:
    """A lat/lon grid of GITM blocks of ``nlons x nlats`` cells."""

    nblocks_lon: int
    nblocks_lat: int
    nalts: int
    nlons: int = 9
    nlats: int = 9
    stretched: bool = False   # geometric dz 1.5 -> 15 km from 100 km
    alt_min: float = 100.0e3
    alt_max: float = 600.0e3

    @property
    def ncols(self) -> int:
        return self.nblocks_lon * self.nblocks_lat * self.nlons * self.nlats


CONFIGS = {
    # 2x2 blocks of 9x9, nAlts 50, uniform 100-600 km
    "cpu_324x50": GridSpec(2, 2, 50),
    # 5 deg x 5 deg: 8 x 4 blocks of 9x9 = 2592 columns, nAlts 50
    "5deg_2592x50": GridSpec(8, 4, 50),
    # 1 deg x 1 deg: 360 x 180 = 40 x 20 blocks of 9x9, nAlts 100 stretched
    "1deg_64800x100": GridSpec(40, 20, 100, stretched=True),
}


def altitudes(spec: GridSpec) -> np.ndarray:
    """Altitude of levels -1..nAlts+2 (metres)."""
    na = spec.nalts
    k = np.arange(-1, na + 3, dtype=np.float64)
    if not spec.stretched:
        dz = (spec.alt_max - spec.alt_min) / na
        return spec.alt_min + (k - 0.5) * dz
    # geometric spacing: dz(1) = 1.5 km growing to dz(nAlts-1) = 15 km
    ratio = 10.0 ** (1.0 / (na - 2))
    dz = 1.5e3 * ratio ** np.arange(na + 3, dtype=np.float64)
    alt = np.empty(na + 4)
    alt[2] = spec.alt_min + 0.5 * dz[0]          # level 1
    for j in range(3, na + 4):
        alt[j] = alt[j - 1] + dz[j - 2]
    alt[1] = alt[2] - dz[0]
    alt[0] = alt[1] - dz[0]
    return alt


def make_columns(spec: GridSpec, seed: int = 1, *, t0: float = 200.0,
                 tinf=1000.0, shape_s: float = 0.02e-3, sigma_t: float = math.log(1.05),
                 w_sigma: float = 5.0, uv_sigma: float = 50.0,
                 tidal: bool | None = None, g0: float = 9.5,
                 eddy_coef: float = 50.0, eddy_p0: float = 0.01,
                 eddy_p1: float = 0.005, ions: bool = False) -> dict:
    """Seeded synthetic columns in the column-contiguous layout.

    ``tinf`` may be a scalar or an array with one exospheric temperature per
    column.  With ``tidal`` (default: on for stretched grids) the horizontal
    winds are three seeded zonal harmonics of 40 m/s amplitude instead of
    white noise.
    """
    if tidal is None:
        tidal = spec.stretched
    na = spec.nalts
    L = na + 4
    nc = spec.ncols
    mass = species_masses()

    nlon_tot = spec.nblocks_lon * spec.nlons
    nlat_tot = spec.nblocks_lat * spec.nlats
    # column c -> (iLon, iLat, block) with lon fastest inside a block
    c = np.arange(nc)
    ilon = c % spec.nlons
    ilat = (c // spec.nlons) % spec.nlats
    blk = c // (spec.nlons * spec.nlats)
    blon = blk % spec.nblocks_lon
    blat = blk // spec.nblocks_lon
    glon = blon * spec.nlons + ilon
    glat = blat * spec.nlats + ilat
    lat = -0.5 * np.pi + (glat + 0.5) * np.pi / nlat_tot
    lon = (glon + 0.5) * 2 * np.pi / nlon_tot

    alt1 = altitudes(spec)
    alt = np.broadcast_to(alt1, (nc, L)).copy()
    dalt = np.empty_like(alt)
    dalt[:, 1:L - 1] = 0.5 * (alt[:, 2:L] - alt[:, 0:L - 2])
    dalt[:, 0] = dalt[:, 1]
    dalt[:, L - 1] = dalt[:, L - 2]
    radial = R_EARTH + alt
    invr = 1 / radial
    grav = -g0 * (R_EARTH * invr) ** 2

    tinf_c = np.broadcast_to(np.asarray(tinf, dtype=np.float64), (nc,))[:, None]
    # Bates profile above 100 km, isothermal below it
    tbase = tinf_c - (tinf_c - t0) * np.exp(-shape_s * np.maximum(alt - spec.alt_min, 0.0))
    tk = tbase * np.exp(sigma_t * normal(seed, 1, (nc, L)))

    # hydrostatic species densities from the level nearest 100 km
    nds = np.empty((nc, NSPECIES, L))
    k100 = 1  # level 0 sits just below 100 km; integrate up and down from it
    for s in range(NSPECIES):
        n = np.empty((nc, L))
        n[:, k100] = BASE_DENSITY[s]
        for j in range(k100 + 1, L):
            gm = -0.5 * (grav[:, j] + grav[:, j - 1])
            tm = 0.5 * (tk[:, j] + tk[:, j - 1])
            n[:, j] = n[:, j - 1] * (tk[:, j - 1] / tk[:, j]) * \
                np.exp(-mass[s] * gm * (alt[:, j] - alt[:, j - 1]) / (KB * tm))
        for j in range(k100 - 1, -1, -1):
            gm = -0.5 * (grav[:, j] + grav[:, j + 1])
            tm = 0.5 * (tk[:, j] + tk[:, j + 1])
            n[:, j] = n[:, j + 1] * (tk[:, j + 1] / tk[:, j]) * \
                np.exp(mass[s] * gm * (alt[:, j + 1] - alt[:, j]) / (KB * tm))
        nds[:, s, :] = n

    vvel = w_sigma * normal(seed, 2, (nc, NSPECIES, L))
    vel = np.empty((nc, 3, L))
    if tidal:
        amp = 40.0
        ph = 2 * np.pi * _uniform(seed, 7, 6)
        ue = np.zeros((nc, L))
        vn = np.zeros((nc, L))
        vert = np.sin(np.pi * (alt - alt1[0]) / (alt1[-1] - alt1[0]))
        for m in range(3):
            ue += amp / 3 * np.cos((m + 1) * lon + ph[m])[:, None] * vert
            vn += amp / 3 * np.sin((m + 1) * lon + ph[m + 3])[:, None] * vert
        vel[:, 0, :] = ue
        vel[:, 1, :] = vn
    else:
        vel[:, 0:2, :] = uv_sigma * normal(seed, 3, (nc, 2, L))
    # bulk vertical wind: mass-weighted species winds
    rho = np.einsum("s,csl->cl", mass, nds)
    vel[:, 2, :] = np.einsum("s,csl->cl", mass, nds * vvel) / rho

    ntot = nds.sum(axis=1)
    mmm = np.zeros((nc, L))
    for s in range(NSPECIES):                       # calc_rates.Earth.f90:36-40
        mmm = mmm + mass[s] * nds[:, s, :] / (ntot + 1.0)
    tempunit = mmm / KB                             # :48
    temp = tk / tempunit                            # normalised temperature
    gam = np.zeros((nc, L))
    for s in range(NSPECIES):                       # calc_pressure.f90:56-86
        gam = gam + nds[:, s, :] / (VIBRATION[s] - 2)
    gam = gam * 2.0 / ntot + 1
    cmax = np.abs(vvel).max(axis=1) + np.sqrt(gam * temp)   # calc_timestep.f90:114-118
    press = ntot * KB * tk
    eddy = np.where(press > eddy_p0, eddy_coef,
                    np.where(press > eddy_p1,
                             eddy_coef * (press - eddy_p1) / (eddy_p0 - eddy_p1), 0.0))

    out = dict(nds=nds, vvel=vvel, vel=vel, temp=temp, rho=rho,
               tempunit=tempunit, mmm=mmm, gamma=gam, cmax=cmax, eddy=eddy,
               alt=alt, grav=grav, invr=invr, dalt=dalt,
               coslat=np.cos(lat), lat=lat, lon=lon)
    if ions:
        # O+ and NO+ Chapman-like layers, other ions small, electrons = sum
        ion = np.empty((nc, NIONS, L))
        zkm = alt / 1e3
        op = 1e12 * np.exp(1 - (zkm - 300) / 60 - np.exp(-(zkm - 300) / 60))
        nop = 1e11 * np.exp(1 - (zkm - 120) / 20 - np.exp(-(zkm - 120) / 20))
        ion[:, 0] = op * np.exp(0.05 * normal(seed, 11, (nc, L))) + 1e3
        ion[:, 1] = nop * np.exp(0.05 * normal(seed, 12, (nc, L))) + 1e3
        for i in range(2, NIONS - 1):
            ion[:, i] = 1e6 + 1e3 * i
        ion[:, NIONS - 1] = ion[:, :NIONS - 1].sum(axis=1)
        out["ions"] = ion
        out["ivel"] = 20.0 * normal(seed, 13, (nc, 3, L))
        out["sza"] = np.arccos(np.clip(np.cos(lat) * np.cos(lon - np.pi), -1, 1))
        out["mlat"] = np.degrees(lat)
    return out


@dataclasses.dataclass
class StepParams:
    """Run parameters of the vertical step (ModInputs defaults unless noted)."""

    dt: float = 2.0
    omega_body: float = 2.0 * 3.141592653589793 / (24.0 * 3600.0)  # ModInputs.f90:285
    max_vert_vel: float = 1000.0       # ModInputs.f90:224
    beta_limiter: float = 2.0          # ModInputs.f90:104
    use_neutral_friction: bool = True  # ModInputs.f90:240
    use_coriolis: bool = True          # ModInputs.f90:235
    use_ion_advection: bool = False    # reference default T (ModInputs.f90:311)
    use_improved_ion_advection: bool = True  # ModInputs.f90:259
    use_nighttime_ion_bcs: bool = True       # ModInputs.f90:260
    use_msis: bool = False             # ModInputs.f90:190 (needs cols["msis"])
    min_tec: float = 2.0               # ModInputs.f90:261
    vert_tau: float = 1.0e9            # ModEarth.f90:325
    use_gswm_tides: bool = False       # ModInputs.f90:195 (needs cols["tides"])
    use_waccm_tides: bool = False      # ModInputs.f90:196 (needs cols["tides"])
    use_damping: bool = False          # ModInputs.f90:221 (Rusanov only)


# ---------------------------------------------------------------------------
# ensembles of members with grids of their own
# ---------------------------------------------------------------------------
ENSEMBLE_NALTS = (50, 60, 70, 80, 90, 100, 110, 120)
# 2592, 10368 and 25920 columns: 8x4, 16x8 and 20x16 blocks of 9x9
ENSEMBLE_BLOCKS = ((8, 4), (16, 8), (20, 16))


@dataclasses.dataclass
class Member:
    """One ensemble member: its grid (uniform 100-600 km levels), exospheric
    temperature and seed (the seed also draws its +-5% perturbations)."""

    spec: GridSpec
    tinf: float
    seed: int


def ensemble(n_members: int = 64, seed: int = 4, nalts=ENSEMBLE_NALTS,
             blocks=ENSEMBLE_BLOCKS, tinf_range=(700.0, 1400.0)) -> list:
    """Members with nAlts drawn uniformly from ``nalts``, block layouts from
    ``blocks`` and the exospheric temperature uniform in ``tinf_range``,
    from counter-based draws of ``seed``."""
    ua = _uniform(seed, 101, n_members)
    ub = _uniform(seed, 102, n_members)
    ut = _uniform(seed, 103, n_members)
    out = []
    for m in range(n_members):
        na = nalts[min(int(ua[m] * len(nalts)), len(nalts) - 1)]
        bl = blocks[min(int(ub[m] * len(blocks)), len(blocks) - 1)]
        tinf = tinf_range[0] + (tinf_range[1] - tinf_range[0]) * ut[m]
        out.append(Member(GridSpec(bl[0], bl[1], na), float(tinf), seed * 1000 + m))
    return out


def member_columns(member: Member, ncols: int | None = None, ions: bool = False) -> dict:
    """The member's columns (all of them, or its first ``ncols``)."""
    spec = member.spec
    if ncols is not None and ncols < spec.ncols:
        # the first whole blocks that hold ncols columns, then cut
        per = spec.nlons * spec.nlats
        nb = -(-ncols // per)
        spec = dataclasses.replace(spec, nblocks_lon=nb, nblocks_lat=1)
    cols = make_columns(spec, seed=member.seed, tinf=member.tinf, ions=ions)
    if ncols is not None and ncols < spec.ncols:
        cols = {k: v[:ncols] for k, v in cols.items()}
    return cols


def msis_bc_values(cols: dict, seed: int = 1, dlog: float = 0.05, dtemp: float = 0.02,
                   wind: float = 20.0) -> np.ndarray:
    """Stand-in values of what msis_bcs returns for each column's level 0
    ([column][9]: LogNS(0,1:6), Temp(0) in K, east and north wind), drawn
    around the column's own level 0 (seeded): log densities +-dlog, the
    temperature within +-dtemp relative, winds N(0, wind) m/s."""
    nc = cols["temp"].shape[0]
    out = np.empty((nc, 9))
    z = normal(seed, 301, (nc, 9))
    out[:, :6] = np.log(cols["nds"][:, :, 1]) + dlog * z[:, :6]
    out[:, 6] = cols["temp"][:, 1] * cols["tempunit"][:, 1] * (1.0 + dtemp * z[:, 6])
    out[:, 7:] = wind * z[:, 7:]
    return out


def tides_values(cols: dict, seed: int = 1, wind: float = 15.0, dtemp: float = 8.0,
                 absolute: bool = False) -> np.ndarray:
    """Stand-in values of GITM's TidesEast, TidesNorth and TidesTemp
    (tides.f90) at each column's levels -1 and 0 ([column][east, north,
    temp][level -1, 0]), seeded: winds N(0, wind) m/s; the temperature a
    perturbation N(0, dtemp) K (GSWM's, added to Temp), or with ``absolute``
    the column's level-0 temperature in K plus that perturbation (WACCM's,
    which replaces it)."""
    nc = cols["temp"].shape[0]
    z = normal(seed, 303, (nc, 3, 2))
    out = np.empty((nc, 3, 2))
    out[:, :2, :] = wind * z[:, :2, :]
    out[:, 2, :] = dtemp * z[:, 2, :]
    if absolute:
        t0 = cols["temp"][:, 0:2] * cols["tempunit"][:, 0:2]
        out[:, 2, :] += t0
    return out


def source_fields(cols: dict, seed: int = 1, heating_peak: float = 150.0e3,
                  heating_width: float = 20.0e3, heating_wkg: float = 1.0e-5) -> dict:
    """Coefficients of the sources step for ``cols`` (synthetic stand-ins of
    GITM's own, seeded): cp as calc_pressure forms it (calc_pressure.f90
    :58-83); KappaTemp = 3.6e-4 T^0.75 W/(m K) and ViscCoef = 1.3e-5
    (T/300)^0.69 kg/(m s), each species' ViscCoefS its density share of
    ViscCoef, all within +-5% (seeded); RadialDistance = R + Altitude (the
    radius the inverse radius of make_columns was formed from); a Gaussian
    heating of
    ``heating_wkg`` W/kg at ``heating_peak`` as the Dt bracket of
    add_sources in Temperature units per second (Q / (cp TempUnit))."""
    nds, temp, tu = cols["nds"], cols["temp"], cols["tempunit"]
    nc, L = temp.shape
    mass = species_masses()
    ntot = nds.sum(axis=1)
    c = np.zeros((nc, L))
    for s in range(NSPECIES):
        c = c + (VIBRATION[s] - 2) * nds[:, s, :] * (KB / mass[s])
    cp = c / (2.0 * ntot)
    tk = temp * tu
    jit = np.exp(0.05 * normal(seed, 401, (nc, 2, L)))
    kappa = 3.6e-4 * tk ** 0.75 * jit[:, 0]
    visc = 1.3e-5 * (tk / 300.0) ** 0.69 * jit[:, 1]
    visc_s = visc[:, None, :] * nds / ntot[:, None, :]
    raddist = R_EARTH + cols["alt"]
    q = heating_wkg * np.exp(-((cols["alt"] - heating_peak) / heating_width) ** 2)
    return dict(cp=cp, kappa_temp=kappa, visc_coef=visc, visc_coef_s=visc_s,
                raddist=raddist, heating=q / (cp * tu))


# --------------------------------------------------------------------------
# Block grids with halo cells, for the horizontal step.

HORIZ_STATE = ("rho", "temp", "vel", "vvel", "nds")
HORIZ_GEOMETRY = ("inv_dlon_gb", "inv_dlon_fb", "dlon_fb", "inv_dlat_gb", "inv_dlat_fb",
                  "dlat_fb", "invr", "raddist")


@dataclasses.dataclass(frozen=True)
class BlockGrid:
    """A global lat/lon grid of ``nblocks_lon x nblocks_lat`` blocks of
    ``nlons x nlats`` cells, uniform altitudes from 100 km."""

    nblocks_lon: int
    nblocks_lat: int
    nalts: int
    nlons: int = 9
    nlats: int = 9
    alt_min: float = 100.0e3
    dz: float = 5.0e3

    @property
    def nblocks(self) -> int:
        return self.nblocks_lon * self.nblocks_lat

    @property
    def ncells(self) -> int:
        return self.nblocks * self.nlons * self.nlats


BLOCK_CONFIGS = {
    # 1 degree: 360 x 180 cells in 9 x 9 blocks, 100 levels
    "1deg_blocks_800x100": BlockGrid(40, 20, 100),
    # 5 degrees: 72 x 36 cells, 50 levels
    "5deg_blocks_32x50": BlockGrid(8, 4, 50),
}


def block_grid(spec: BlockGrid, seed: int = 1, ions: bool = False, blocks=None,
               topography: bool = False) -> dict:
    """Block-layout state, geometry and cMax of a smooth synthetic
    atmosphere.  Every per-cell field is [block][component][level][lat][lon]
    with levels -1..nAlts+2 and lat/lon -1..n+2 (halo cells hold the
    neighbouring blocks' values, as after GITM's message exchange; cells
    past a pole continue the analytic fields).  The geometry follows
    initialize.f90:185-267 (cell sizes half way between neighbours, face
    distances between centres, cos(latitude) floored at 0.01, tangent
    capped at 100); cMax(east, north) = |v| + sqrt(gamma T) on cells 0..n+1
    as calc_timestep_horizontal (calc_timestep.f90:34-41), up as the
    vertical CFL speed.  ``blocks``: the block numbers to generate (all by
    default; block b is at longitude b % nblocks_lon, latitude
    b // nblocks_lon).  ``topography``: the levels follow a terrain of a few
    km that fades with height, Altitude_GB = alt + h(lon, lat) exp(-(alt -
    alt_min) / 200 km); adds "dalt" (dAlt_GB, half the distance between the
    neighbouring levels) and "dalt_dlon", "dalt_dlat" (dAltDLon_CB,
    dAltDLat_CB: the slopes of the level surfaces per metre east and north,
    initialize.f90:294-300), all shaped as a state field."""
    ids = np.arange(spec.nblocks) if blocks is None else np.asarray(blocks, dtype=np.int64)
    nb, nlo, nla, nA = len(ids), spec.nlons, spec.nlats, spec.nalts
    L, NLo, NLg = nA + 4, nlo + 4, nla + 4
    dlon = 2.0 * np.pi / (spec.nblocks_lon * nlo)
    dlat = np.pi / (spec.nblocks_lat * nla)
    bi = ids % spec.nblocks_lon
    bj = ids // spec.nblocks_lon
    ii = np.arange(-1, nlo + 3)   # lon index -1..nLons+2
    jj = np.arange(-1, nla + 3)
    lon = (bi[:, None] * nlo + ii[None, :] - 0.5) * dlon           # [nb][NLo]
    lat = -0.5 * np.pi + (bj[:, None] * nla + jj[None, :] - 0.5) * dlat  # [nb][NLg]
    alt = spec.alt_min + (np.arange(L) - 1.5) * spec.dz
    r = R_EARTH + alt
    shp = (nb, L, NLg, NLo)
    LON = np.broadcast_to(lon[:, None, None, :], shp)
    LAT = np.broadcast_to(lat[:, None, :, None], shp)
    Z = np.broadcast_to(alt[None, :, None, None], shp)
    R = np.broadcast_to(r[None, :, None, None], shp)
    ph = 0.7 * seed
    T = (1000.0 - 800.0 * np.exp(-(Z - 100.0e3) / 50.0e3)) * (
        1.0 + 0.05 * np.cos(LAT) * np.sin(LON + ph) + 0.02 * np.sin(3 * LAT + ph))
    m = species_masses()
    nds = np.empty((nb, NSPECIES) + shp[1:])
    for s in range(NSPECIES):
        H = KB * 600.0 / (m[s] * 9.5)
        nds[:, s] = BASE_DENSITY[s] * np.exp(-(Z - 100.0e3) / H) * (
            1.0 + 0.1 * np.sin(2 * LON + s + ph) * np.cos(LAT) + 0.03 * np.cos(LAT * (s + 1)))
    rho = np.einsum("s,bs...->b...", m, nds)
    ntot = nds.sum(axis=1)
    mmm = rho / ntot
    temp = T / (mmm / KB)                       # Temperature = T / TempUnit
    vel = np.empty((nb, 3) + shp[1:])
    vel[:, 0] = 80.0 * np.cos(LAT) * np.sin(LON + 2 * LAT + ph) + 20.0 * np.sin(3 * LON)
    vel[:, 1] = 60.0 * np.sin(2 * LAT) * np.cos(LON - ph) + 10.0 * np.cos(2 * LON + LAT)
    vel[:, 2] = 3.0 * np.sin(LON + LAT + ph) * (Z / 400.0e3)
    vvel = np.empty((nb, NSPECIES) + shp[1:])
    for s in range(NSPECIES):
        vvel[:, s] = vel[:, 2] + (1.0 + 0.3 * s) * np.cos(LON * (s % 3 + 1) + LAT + ph)
    gamma = 1.0 + 2.0 / ntot * sum(nds[:, s] / (VIBRATION[s] - 2.0) for s in range(NSPECIES))
    out = {"rho": rho, "temp": temp, "vel": vel, "vvel": vvel, "nds": nds, "gamma": gamma}
    if ions:
        base = (2e11, 1e11, 5e10, 1e9, 1e10, 1e8, 1e8, 1e9, 1e8)
        idens = np.empty((nb, NIONS) + shp[1:])
        for s in range(NIONS - 1):
            idens[:, s] = base[s] * np.exp(-((Z - 300.0e3) / 150.0e3) ** 2) * (
                1.0 + 0.2 * np.cos(LON + s * 0.3 + ph) * np.cos(LAT)) + 1.0e3
        idens[:, NIONS - 1] = idens[:, :NIONS - 1].sum(axis=1)
        ivel = np.empty((nb, 3) + shp[1:])
        ivel[:, 0] = 150.0 * np.sin(LON + ph) * np.cos(LAT)
        ivel[:, 1] = 100.0 * np.cos(2 * LON) * np.sin(LAT)
        ivel[:, 2] = 10.0 * np.sin(LAT + ph)
        ipar = np.empty((nb, 3) + shp[1:])  # the field-aligned part
        ipar[:, 0] = 40.0 * np.cos(LON + ph) * np.sin(2 * LAT)
        ipar[:, 1] = 60.0 * np.sin(LAT) * np.cos(LON)
        ipar[:, 2] = 30.0 * np.sin(LAT + ph) * (Z / 400.0e3)
        out["ions"], out["ivel"], out["ivel_par"] = idens, ivel, ipar
        # UAM_Gradient's coefficients (initialize.f90:268-292), [block][M, 0, P][cell]
        for key, ang, n in (("grad_lon", lon, nlo), ("grad_lat", lat, nla)):
            dm = ang[:, 2:n + 2] - ang[:, 1:n + 1]   # cells 1..n minus cells 0..n-1
            dp = ang[:, 3:n + 3] - ang[:, 2:n + 2]
            r2 = (dm / dp) ** 2
            inv = 1.0 / (r2 * dp + dm)
            out[key] = np.stack([-inv, inv * (1 - r2), inv * r2], axis=1)
    # geometry (initialize.f90:185-267)
    cosl = np.cos(lat)
    fl = np.maximum(np.abs(cosl), 0.01)[:, None, :, None]
    lonp = np.concatenate([lon[:, :1] - dlon, lon, lon[:, -1:] + dlon], axis=1)  # -2..n+3
    latp = np.concatenate([lat[:, :1] - dlat, lat, lat[:, -1:] + dlat], axis=1)
    Rm_lon = np.concatenate([R[..., :1], R[..., :-1]], axis=3)  # r at lon-1
    Rm_lat = np.concatenate([R[:, :, :1], R[:, :, :-1]], axis=2)
    dlat_gb = 0.5 * (latp[:, 2:] - latp[:, :-2])[:, None, :, None] * R
    dlat_fb = (latp[:, 1:-1] - latp[:, :-2])[:, None, :, None] * (0.5 * (R + Rm_lat))
    dlon_gb = 0.5 * (lonp[:, 2:] - lonp[:, :-2])[:, None, None, :] * R * fl
    dlon_fb = (lonp[:, 1:-1] - lonp[:, :-2])[:, None, None, :] * (0.5 * (R + Rm_lon)) * fl
    out.update({"inv_dlon_gb": 1.0 / dlon_gb, "inv_dlon_fb": 1.0 / dlon_fb,
                "dlon_fb": np.ascontiguousarray(dlon_fb),
                "inv_dlat_gb": 1.0 / dlat_gb, "inv_dlat_fb": 1.0 / dlat_fb,
                "dlat_fb": np.ascontiguousarray(dlat_fb),
                "invr": 1.0 / np.ascontiguousarray(R), "raddist": np.ascontiguousarray(R),
                "latitude": np.ascontiguousarray(lat),
                "tan_lat": np.minimum(np.abs(np.tan(lat)), 100.0) * np.sign(lat),
                "cos_lat": np.ascontiguousarray(cosl)})
    if topography:
        fade = np.exp(-(Z - spec.alt_min) / 200.0e3)
        h = 3000.0 * (1.0 + np.sin(2 * LON + ph) * np.cos(LAT)) + 800.0 * np.cos(3 * LAT)
        dh_dlon = 6000.0 * np.cos(2 * LON + ph) * np.cos(LAT)
        dh_dlat = -3000.0 * np.sin(2 * LON + ph) * np.sin(LAT) - 2400.0 * np.sin(3 * LAT)
        A = Z + h * fade
        dalt = np.empty(shp)
        dalt[:, 1:-1] = 0.5 * (A[:, 2:] - A[:, :-2])
        dalt[:, 0], dalt[:, -1] = dalt[:, 1], dalt[:, -2]
        out["dalt"] = dalt
        out["dalt_dlon"] = dh_dlon * fade / (R * np.maximum(np.abs(np.cos(LAT)), 0.01))
        out["dalt_dlat"] = dh_dlat * fade / R
    # cMax_GDB [nb][3][0..nAlts+1][0..nLats+1][0..nLons+1] (iEast_, iNorth_, iUp_)
    cs = np.sqrt(gamma * temp)[:, 1:nA + 3, 1:nla + 3, 1:nlo + 3]
    cmax = np.empty((nb, 3, nA + 2, nla + 2, nlo + 2))
    cmax[:, 0] = np.abs(vel[:, 0, 1:nA + 3, 1:nla + 3, 1:nlo + 3]) + cs
    cmax[:, 1] = np.abs(vel[:, 1, 1:nA + 3, 1:nla + 3, 1:nlo + 3]) + cs
    cmax[:, 2] = np.abs(vvel[:, :, 1:nA + 3, 1:nla + 3, 1:nlo + 3]).max(axis=1) + cs
    out["cmax"] = cmax
    return {k: np.ascontiguousarray(v, dtype=np.float64) for k, v in out.items()}


def horizontal_dt(grid: dict, cfl: float = 0.25) -> float:
    """The horizontal CFL step of calc_timestep_horizontal
    (calc_timestep.f90:43-50): cfl / max(cMax_E / dLon + cMax_N / dLat) over
    interior cells and levels 1..nAlts."""
    c = grid["cmax"][:, :, 1:-1, 1:-1, 1:-1]
    igl = grid["inv_dlon_gb"][:, 2:-2, 2:-2, 2:-2]
    iga = grid["inv_dlat_gb"][:, 2:-2, 2:-2, 2:-2]
    return float(cfl / np.max(c[:, 0] * igl + c[:, 1] * iga))
