"""Python binding of the MI355X SPH-DEM library (libsphdem.so, C ABI in
include/sphdem_c.h) plus the seeded initial-condition generators of the
benchmark workloads.

The reference drives its loop from C++ (run.cpp:16-164) and only uses Python to
write parameters.h and launch runs (sediment3D/runAll.py).  This module is the
counterpart used by tests and the benchmark: it builds an `sphdem_params`
(the values a case's customConstants.h fixes at compile time) and the particle
arrays of a setup.cpp-style initial condition, and calls the C ABI through
ctypes.  The library fails loudly when the HIP build is missing: there is no
CPU fallback.
"""
from __future__ import annotations

import ctypes
import math
import os
from dataclasses import dataclass, field

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libsphdem.so")

# particle types: enum iamTypes (particle.h:17)
SPH, SPH_BOUNDARY, IMMOVABLE, BOUNDARY, GHOST, DEM = 0, 1, 2, 3, 4, 11
DELETED = 255

F_VAR_H_CORRECTION = 1 << 0
F_CONST_H = 1 << 1
F_LIQ_DEM = 1 << 2
F_VISC_MORRIS = 1 << 3
F_VISC_ARTIFICIAL = 1 << 4
F_CUSTOM_WALL = 1 << 5
F_NO_DEM_CONTACTS = 1 << 6
F_FIXED_DEM = 1 << 7
F_VISC_MONAGHAN = 1 << 8
F_VISC_CLEARY = 1 << 9
F_DEM_BLOCK = 1 << 10

FP64, MIXED = 0, 1
# return codes of the C ABI (SPHDEM_E_*)
E_ARG, E_HIP, E_STATE, E_CAPACITY = -1, -2, -3, -4

HOOKS_NONE, HOOKS_DAMBREAK, HOOKS_SEDIMENT = 0, 1, 2

FIELD = {
    "r": 0, "v": 1, "v0": 2, "vhat": 3, "vhat0": 4, "f": 5, "fp": 6, "fv": 7,
    "fb": 8, "fdrag": 9, "fvel": 10,
    "dens": 16, "dddt": 17, "h": 18, "mass": 19, "press": 20, "pdr2": 21,
    "omega": 22, "porosity": 23, "shepSum": 24, "u": 25, "dudt": 26,
    "iam": 32, "tag": 33,
}
VECTOR_FIELDS = {k for k, v in FIELD.items() if v < 16}
INT_FIELDS = {"iam", "tag"}

PAIRS_SPH, PAIRS_DEM, PAIRS_SPH2DEM, PAIRS_DEM2SPH = 0, 1, 2, 3


class Params(ctypes.Structure):
    """Mirror of `sphdem_params` (include/sphdem_c.h)."""

    _fields_ = [
        ("flags", ctypes.c_uint32),
        ("precision", ctypes.c_int32),
        ("periodic", ctypes.c_int32 * 3),
        ("reinit_every", ctypes.c_int32),
        ("reinit_at", ctypes.c_int32),
        ("damptime", ctypes.c_int32),
        ("rmin", ctypes.c_double * 3),
        ("rmax", ctypes.c_double * 3),
        ("proc_domain", ctypes.c_double * 6),
        ("pi", ctypes.c_double),
        ("dens", ctypes.c_double),
        ("refd", ctypes.c_double),
        ("gamma", ctypes.c_double),
        ("hfac", ctypes.c_double),
        ("h", ctypes.c_double),
        ("psep", ctypes.c_double),
        ("spsound", ctypes.c_double),
        ("prb", ctypes.c_double),
        ("viscosity", ctypes.c_double),
        ("alpha_artificial", ctypes.c_double),
        ("wcon", ctypes.c_double),
        ("maxdt", ctypes.c_double),
        ("dem_radius", ctypes.c_double),
        ("dem_dens", ctypes.c_double),
        ("dem_vol", ctypes.c_double),
        ("dem_k", ctypes.c_double),
        ("dem_gamma", ctypes.c_double),
        ("dem_search_radius", ctypes.c_double),
        ("coupling_radius", ctypes.c_double),
        ("time_drop", ctypes.c_double),
        ("dem_timestep", ctypes.c_double),
        ("liq_dem_timestep", ctypes.c_double),
        ("porosity_incr", ctypes.c_double),
        ("drag_a", ctypes.c_double),
        ("drag_b", ctypes.c_double),
    ]

    def copy(self) -> "Params":
        p = Params()
        ctypes.pointer(p)[0] = self
        return p


class Output(ctypes.Structure):
    """Mirror of `sphdem_output` (calcOutputVars diagnostics)."""

    _fields_ = [("linMom", ctypes.c_double * 3)] + [
        (k, ctypes.c_double) for k in (
            "angMom", "eKE", "eViscF", "edViscFdt", "eViscB", "eBForce", "eFF", "edFFdt",
            "eElast", "edElastdt", "eElastExact", "eTotal", "maxV", "maxF", "maxFF", "rmsFF",
            "aveDens", "varDens", "aveDensFromMass")] + [("n", ctypes.c_int64),
                                                         ("nSph", ctypes.c_int64)]


class Globals(ctypes.Structure):
    """Mirror of `sphdem_globals`."""

    _fields_ = [
        ("time", ctypes.c_double),
        ("dt", ctypes.c_double),
        ("newDt", ctypes.c_double),
        ("maxdt", ctypes.c_double),
        ("sphStep", ctypes.c_int64),
        ("custom", ctypes.c_double * 9),
    ]


# ----------------------------------------------------------------------------
# seeded initial conditions
# ----------------------------------------------------------------------------
PI_REF = 3.14159265  # PI of the reference cases (sediment3D/customConstants.h:11)
BIG_DT = 1.0e10      # maxdt = "infinity": the dt rule alone sets the step


def splitmix64(seed: int, n: int) -> np.ndarray:
    """n uniform doubles in [0,1) from splitmix64 seeded with `seed` (the
    counter-based generator every initial condition here uses)."""
    mask = (1 << 64) - 1
    out = np.empty(n, dtype=np.float64)
    x = seed & mask
    # vectorised: state_k = seed + k*gamma
    gamma = 0x9E3779B97F4A7C15
    k = np.arange(1, n + 1, dtype=np.uint64)
    with np.errstate(over="ignore"):
        z = (np.uint64(x) + k * np.uint64(gamma))
        z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
        z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
        z = z ^ (z >> np.uint64(31))
    out[:] = (z >> np.uint64(11)).astype(np.float64) * (1.0 / (1 << 53))
    return out


@dataclass
class Case:
    """A complete initial condition: constants + particle arrays."""

    name: str
    params: Params
    r: np.ndarray
    v: np.ndarray
    vhat: np.ndarray
    mass: np.ndarray
    h: np.ndarray
    dens: np.ndarray
    iam: np.ndarray
    hookset: int
    info: dict = field(default_factory=dict)

    @property
    def n(self) -> int:
        return int(self.r.shape[0])

    def counts(self) -> dict:
        return {
            "fluid": int(np.sum(self.iam == SPH)),
            "wall": int(np.sum(self.iam == SPH_BOUNDARY)),
            "dem": int(np.sum(self.iam == DEM)),
        }


def _base_params(flags: int, rmin, rmax, periodic, dx: float, spsound: float,
                 viscosity: float, gamma: float = 7.0) -> Params:
    P = Params()
    P.flags = flags
    P.precision = FP64
    for i in range(3):
        P.periodic[i] = int(periodic[i])
        P.rmin[i] = rmin[i]
        P.rmax[i] = rmax[i]
        L = rmax[i] - rmin[i]
        if periodic[i]:
            P.proc_domain[2 * i] = rmin[i]
            P.proc_domain[2 * i + 1] = rmax[i]
        else:  # splitDomain, misc.cpp:726-737
            P.proc_domain[2 * i] = rmin[i] - 0.25 * L
            P.proc_domain[2 * i + 1] = rmax[i] + 0.25 * L
    P.reinit_every = 2 ** 31 - 1
    P.reinit_at = -1
    P.damptime = 0
    P.pi = PI_REF
    P.dens = 1000.0
    P.refd = 1000.0
    P.gamma = gamma
    P.hfac = 1.5
    P.psep = dx
    P.h = P.hfac * dx
    P.spsound = spsound
    # PRB = (REFD/DENS)^(GAMMA-1) c^2 REFD/GAMMA (sediment3D/customConstants.h:116)
    P.prb = math.pow(P.refd / P.dens, gamma - 1.0) * spsound ** 2 * P.refd / gamma
    P.viscosity = viscosity
    P.alpha_artificial = 0.1
    P.wcon = 21.0 / (256.0 * PI_REF)
    P.maxdt = BIG_DT
    P.drag_a = 3.7
    P.drag_b = 0.65
    return P


def _set_dem_constants(P: Params, radius: float, dem_dens: float, fluid_dt: float,
                       restitution: float = 0.9, coupling_radius: float | None = None):
    """Linear spring-dashpot constants with restitution e and DEM_TIMESTEP =
    t_c/50 (demConstants.h:6) equal to `fluid_dt`."""
    P.dem_radius = radius
    P.dem_dens = dem_dens
    P.dem_vol = (4.0 / 3.0) * PI_REF * radius ** 3
    P.dem_search_radius = radius
    mass = P.dem_vol * dem_dens
    m_red = 0.5 * mass
    t_c = 50.0 * fluid_dt
    beta = -math.log(restitution) / t_c
    omega0_sq = (math.pi / t_c) ** 2 + beta ** 2
    P.dem_k = m_red * omega0_sq
    P.dem_gamma = 2.0 * m_red * beta
    P.dem_timestep = (1.0 / 50.0) * PI_REF / math.sqrt(
        P.dem_k / m_red - math.pow(0.5 * P.dem_gamma / m_red, 2))
    P.liq_dem_timestep = (1.0 / 5.0) * P.dem_vol * dem_dens / (
        6.0 * PI_REF * P.viscosity * P.dens * radius)
    P.coupling_radius = coupling_radius if coupling_radius is not None else 2.0 * P.h
    P.time_drop = 0.0
    P.porosity_incr = (4.0 / 3.0) * PI_REF * math.pow(radius, 3)


def _lattice(ni, nj, nk, dx, origin=(0.0, 0.0, 0.0)):
    i, j, k = np.meshgrid(np.arange(ni), np.arange(nj), np.arange(nk), indexing="ij")
    pts = np.stack([(i + 0.5) * dx + origin[0], (j + 0.5) * dx + origin[1],
                    (k + 0.5) * dx + origin[2]], axis=-1).reshape(-1, 3)
    return pts


def _jitter(pts: np.ndarray, dx: float, seed: int, frac: float = 0.01) -> np.ndarray:
    u = splitmix64(seed, pts.size).reshape(pts.shape)
    return pts + (2.0 * u - 1.0) * frac * dx


def _place_dem(n: int, lo, hi, radius: float, seed: int, periodic, box) -> np.ndarray:
    """Rejection sampling of n non-overlapping spheres (|dr| > 2R, minimum image
    on periodic axes) uniformly in [lo, hi)."""
    cell = 2.0 * radius
    L = [float(x) for x in box]
    ncell = [max(1, int(math.floor(L[a] / cell))) for a in range(3)]
    lo = [float(x) for x in lo]
    span = [float(hi[a]) - lo[a] for a in range(3)]
    r2min = 4.0 * radius * radius

    def key(p):
        c = []
        for a in range(3):
            ci = int(math.floor(p[a] / cell))
            if periodic[a]:
                ci %= ncell[a]  # the last cell absorbs the remainder of L
            c.append(ci)
        return c

    grid: dict = {}
    out = np.empty((n, 3))
    got = 0
    batch = 3 * max(1024, 4 * n)
    u = splitmix64(seed, batch).tolist()
    pos = 0
    refill = 0
    while got < n:
        if pos + 3 > len(u):
            refill += 1
            u = splitmix64(seed + 7919 * refill, batch).tolist()
            pos = 0
        p = [lo[a] + u[pos + a] * span[a] for a in range(3)]
        pos += 3
        c = key(p)
        ok = True
        for di in (-1, 0, 1):
            for dj in (-1, 0, 1):
                for dk in (-1, 0, 1):
                    cc = [c[0] + di, c[1] + dj, c[2] + dk]
                    for a in range(3):
                        if periodic[a]:
                            cc[a] %= ncell[a]
                    for q in grid.get((cc[0], cc[1], cc[2]), ()):
                        d2 = 0.0
                        for a in range(3):
                            d = p[a] - out[q, a]
                            if periodic[a]:
                                d -= L[a] * round(d / L[a])
                            d2 += d * d
                        if d2 <= r2min:
                            ok = False
                            break
                    if not ok:
                        break
                if not ok:
                    break
            if not ok:
                break
        if not ok:
            continue
        out[got] = p
        grid.setdefault((c[0], c[1], c[2]), []).append(got)
        got += 1
    return out


def _wall_shell(lo_idx, hi_idx, inner_lo, inner_hi, dx):
    """Lattice points (i+0.5)dx with lo_idx <= i < hi_idx outside the inner box."""
    i, j, k = np.meshgrid(np.arange(lo_idx[0], hi_idx[0]), np.arange(lo_idx[1], hi_idx[1]),
                          np.arange(lo_idx[2], hi_idx[2]), indexing="ij")
    inside = ((i >= inner_lo[0]) & (i < inner_hi[0]) & (j >= inner_lo[1]) & (j < inner_hi[1])
              & (k >= inner_lo[2]) & (k < inner_hi[2]))
    sel = ~inside
    pts = np.stack([(i[sel] + 0.5) * dx, (j[sel] + 0.5) * dx, (k[sel] + 0.5) * dx], axis=-1)
    return pts


def _finish(name, P, fluid, walls, dem, hookset, wall_dens, info):
    nf, nw, nd = len(fluid), len(walls), len(dem)
    r = np.concatenate([fluid, walls, dem]).astype(np.float64)
    n = len(r)
    iam = np.concatenate([np.full(nf, SPH), np.full(nw, SPH_BOUNDARY),
                          np.full(nd, DEM)]).astype(np.int32)
    mass = np.empty(n)
    dens = np.empty(n)
    h = np.empty(n)
    m_f = P.psep ** 3 * P.dens  # setup.cpp:80 p.mass = pow(PSEP,NDIM)*DENS
    mass[: nf + nw] = m_f
    dens[:nf] = P.dens
    dens[nf: nf + nw] = wall_dens
    h[: nf + nw] = P.hfac * (m_f / P.dens) ** (1.0 / 3.0)  # setup.cpp:81, before the wall density
    if nd:
        mass[nf + nw:] = P.dem_vol * P.dem_dens
        dens[nf + nw:] = P.dem_dens
        h[nf + nw:] = P.dem_search_radius
    z = np.zeros((n, 3))
    return Case(name, P, r, z.copy(), z.copy(), mass, h, dens, iam, hookset, info)


# ---- files for the C++ host (sph-dem_amd/host): constants and initial conditions
_FLAG_DEFINES = [
    (F_VAR_H_CORRECTION, "VAR_H_CORRECTION"), (F_CONST_H, "CONST_H"), (F_LIQ_DEM, "LIQ_DEM"),
    (F_VISC_MORRIS, "VISC_MORRIS"), (F_VISC_ARTIFICIAL, "VISC_ARTIFICIAL"),
    (F_CUSTOM_WALL, "LIQ_DEM_CUSTOM_WALL_CONTACTS"), (F_NO_DEM_CONTACTS, "NO_DEM_CONTACTS"),
    (F_FIXED_DEM, "FIXED_DEM"), (F_VISC_MONAGHAN, "VISC_MONAGHAN"), (F_VISC_CLEARY, "VISC_CLEARY"),
    (F_DEM_BLOCK, "DEM_BLOCK"),
]


def custom_constants_header(P: Params, title: str, max_particles: int = 1) -> str:
    """A customConstants.h with the names the reference's core headers read,
    holding exactly the doubles of P (17 significant digits round-trip).  It
    defines SPHDEM_DERIVED_CONSTANTS and carries the constants the reference
    derives in its core headers (WCON_WENDLAND, DEM_TIMESTEP, ...) as the
    exact values of P, and MAX_NUM_PARTICLES_PER_CPU = max_particles."""
    def d(x):
        return repr(float(x))

    out = [f"// customConstants.h -- {title}.", "// Generated by sphdem.custom_constants_header().",
           "#pragma once", "", "#define SPHDEM_DERIVED_CONSTANTS", "#define LINEAR"]
    for bit, name in _FLAG_DEFINES:
        if P.flags & bit:
            out.append(f"#define {name}")
    out.append("")
    out.append("const int PERIODIC[3] = {%d, %d, %d};" % tuple(P.periodic))
    out.append("const double RMIN[3] = {%s, %s, %s};" % tuple(d(x) for x in P.rmin))
    out.append("const double RMAX[3] = {%s, %s, %s};" % tuple(d(x) for x in P.rmax))
    ints = [("REINIT_DENS_EVERY_N_STEPS", P.reinit_every), ("REINIT_DENS_AT_N_STEPS", P.reinit_at),
            ("DAMPTIME", P.damptime), ("MAX_NUM_PARTICLES_PER_CPU", max_particles)]
    for name, v in ints:
        out.append(f"const int {name} = {int(v)};")
    dbl = [("PI", P.pi), ("DENS", P.dens), ("REFD", P.refd), ("GAMMA", P.gamma),
           ("HFAC", P.hfac), ("H", P.h), ("PSEP", P.psep), ("SPSOUND", P.spsound),
           ("PRB", P.prb), ("VISCOSITY", P.viscosity), ("ALPHA_ARTIFICIAL", P.alpha_artificial),
           ("WCON_WENDLAND", P.wcon), ("MAXDT", P.maxdt), ("DEM_RADIUS", P.dem_radius),
           ("DEM_DENS", P.dem_dens), ("DEM_VOL", P.dem_vol), ("DEM_K", P.dem_k),
           ("DEM_GAMMA", P.dem_gamma), ("DEM_SEARCH_RADIUS", P.dem_search_radius),
           ("LIQ_DEM_COUPLING_RADIUS", P.coupling_radius), ("TIME_DROP_PARTICLE", P.time_drop),
           ("DEM_TIMESTEP", P.dem_timestep), ("LIQ_DEM_TIMESTEP", P.liq_dem_timestep),
           ("POROSITY_INCR", P.porosity_incr), ("DRAG_A", P.drag_a), ("DRAG_B", P.drag_b)]
    for name, v in dbl:
        out.append(f"const double {name} = {d(v)};")
    return "\n".join(out) + "\n"


IC_MAGIC = b"SPHDEMIC"
STATE_MAGIC = b"SPHDEMST"


def write_ic(case: Case, path: str):
    """Initial condition file the C++ host's run driver reads: magic, int64 n,
    then r, v, vhat (3n doubles each), mass, h, dens (n doubles) and iam (n int32)."""
    with open(path, "wb") as f:
        f.write(IC_MAGIC)
        f.write(np.int64(case.n).tobytes())
        for a in (case.r, case.v, case.vhat):
            f.write(np.ascontiguousarray(a, dtype=np.float64).tobytes())
        for a in (case.mass, case.h, case.dens):
            f.write(np.ascontiguousarray(a, dtype=np.float64).tobytes())
        f.write(np.ascontiguousarray(case.iam, dtype=np.int32).tobytes())


STATE_FIELDS = [("r", 3), ("v", 3), ("vhat", 3), ("f", 3), ("dens", 1), ("h", 1),
                ("press", 1), ("omega", 1), ("porosity", 1)]


def read_state(path: str) -> dict:
    """State file written by the C++ run driver: magic, int64 n, time, dt, then
    the STATE_FIELDS arrays per original index."""
    with open(path, "rb") as f:
        if f.read(8) != STATE_MAGIC:
            raise ValueError(f"{path}: not a state file")
        n = int(np.frombuffer(f.read(8), dtype=np.int64)[0])
        time, dt = np.frombuffer(f.read(16), dtype=np.float64)
        out = {"time": float(time), "dt": float(dt)}
        for name, w in STATE_FIELDS:
            a = np.frombuffer(f.read(8 * w * n), dtype=np.float64)
            out[name] = a.reshape(n, w) if w == 3 else a.copy()
    return out


def dambreak_case(nf: int = 32, dx: float = 0.01, seed: int = 1000, with_dem: int = 0,
                  liq_dem_flags: bool = False) -> Case:
    """3D dam break: an nf^3 fluid block (jittered lattice) in the lower-left
    half of a 2L x L x L box closed by 3 layers of sphBoundary walls; Tait EOS,
    Wendland, Morris + artificial (alpha = 0.1) viscosity, gravity on sph.
    With `with_dem` > 0 and the SPH-DEM flag set it becomes the large
    mixed-precision / adjoint geometry: DEM spheres placed in the upper half of
    the fluid block."""
    Lc = nf * dx
    rmin = (0.0, 0.0, 0.0)
    rmax = (2 * Lc, Lc, Lc)
    c0 = 10.0 * math.sqrt(2.0 * 9.81 * Lc)
    if liq_dem_flags:
        flags = F_VISC_MORRIS | F_VAR_H_CORRECTION | F_LIQ_DEM | F_CUSTOM_WALL
        visc = 8.9e-7
    else:
        flags = F_VISC_MORRIS | F_VISC_ARTIFICIAL
        visc = 1.0e-6
    P = _base_params(flags, rmin, rmax, (0, 0, 0), dx, c0, visc)
    fluid = _jitter(_lattice(nf, nf, nf, dx), dx, seed)
    walls = _wall_shell((-3, -3, -3), (2 * nf + 3, nf + 3, nf + 3), (0, 0, 0),
                        (2 * nf, nf, nf), dx)
    wall_dens = P.refd * math.pow(9.81 * P.refd * Lc / P.prb + 1, 1.0 / 7.0)
    dem = np.zeros((0, 3))
    if liq_dem_flags:
        fluid_dt = 0.6 * P.h / (2.0 * P.spsound)
        _set_dem_constants(P, 0.25 * dx, 2500.0, fluid_dt)
        if with_dem:
            R = P.dem_radius
            dem = _place_dem(with_dem, (R, R, 0.5 * Lc), (Lc - R, Lc - R, Lc - R), R,
                             seed + 1, (0, 0, 0), rmax)
    hook = HOOKS_SEDIMENT if liq_dem_flags else HOOKS_DAMBREAK
    return _finish("dambreak", P, fluid, walls, dem, hook, wall_dens,
                   {"nf": nf, "dx": dx, "seed": seed})


def sediment_case(nf: int = 64, n_dem: int | None = None, seed: int = 1001,
                  jitter: bool = True, nz: int | None = None) -> Case:
    """SPH-DEM sedimentation: an nf^3 fluid lattice (dx = 1/nf) filling the unit
    cube, periodic in x and y, 3 layers of sphBoundary walls below z=0 and above
    z=1, and n_dem (default nf^3/64) DEM spheres of diameter dx/2, density 2500,
    placed without overlap in the upper half.  With `nz` the fluid has nz
    layers and the box is nz*dx high (nf x nf x nz particles).  sediment3D flag set
    (sediment3D/customConstants.h:45-79): WENDLAND, VISC_MORRIS,
    VAR_H_CORRECTION, LIQ_DEM, LINEAR, NO_ANTICLUMPING,
    LIQ_DEM_CUSTOM_WALL_CONTACTS; coupling radius 2h, Di Felice drag,
    restitution 0.9."""
    dx = 1.0 / nf
    lz = 1.0 if nz is None else nz * dx
    nz = nf if nz is None else nz
    rmin = (0.0, 0.0, 0.0)
    rmax = (1.0, 1.0, lz)
    vmax = 2.0 * math.sqrt(2 * 9.81 * (rmax[2] - rmin[2]))  # customConstants.h:97
    c0 = 10.0 * vmax
    flags = F_VISC_MORRIS | F_VAR_H_CORRECTION | F_LIQ_DEM | F_CUSTOM_WALL
    P = _base_params(flags, rmin, rmax, (1, 1, 0), dx, c0, 8.9e-7)
    fluid = _lattice(nf, nf, nz, dx)
    if jitter:
        fluid = _jitter(fluid, dx, seed)
    lo = _lattice(nf, nf, 3, dx, origin=(0.0, 0.0, -3 * dx))
    hi = _lattice(nf, nf, 3, dx, origin=(0.0, 0.0, lz))
    walls = np.concatenate([lo, hi])
    wall_dens = P.refd * math.pow(9.81 * P.refd * (rmax[2] - rmin[2]) / P.prb + 1, 1.0 / 7.0)
    fluid_dt = 0.6 * P.h / (2.0 * P.spsound)
    _set_dem_constants(P, 0.25 * dx, 2500.0, fluid_dt)
    if n_dem is None:
        n_dem = nf * nf * nz // 64
    R = P.dem_radius
    dem = _place_dem(n_dem, (0.0, 0.0, 0.5 * lz), (1.0, 1.0, lz - R), R, seed + 1,
                     (1, 1, 0), rmax) if n_dem else np.zeros((0, 3))
    return _finish("sediment", P, fluid, walls, dem, HOOKS_SEDIMENT, wall_dens,
                   {"nf": nf, "nz": nz, "dx": dx, "seed": seed, "n_dem": n_dem})


def adjoint_case(nf: int = 64, nz: int = 16, n_dem: int = 8192, seed: int = 1004,
                 vscale: float = 0.05, touching: int = 0, floor: int = 0) -> Case:
    """The adjoint workload: the sedimentation geometry (sediment_case: nf x nf x
    nz fluid, dx = 1/nf, periodic in x and y) with n_dem DEM spheres, seeded
    initial velocities uniform in [-vscale, vscale) per component on fluid and
    DEM particles (walls at rest), the drop time moved before t = 0 so that
    no particle is frozen, `touching` DEM pairs moved into contact (centre
    distance 1.9 R) and `floor` DEM spheres moved onto the z = 0 floor
    (centre at 0.8 R), so that the contact terms take part."""
    c = sediment_case(nf=nf, n_dem=n_dem, seed=seed, nz=nz)
    P = c.params
    P.time_drop = -1.0
    R = P.dem_radius
    dem = np.where(c.iam == DEM)[0]
    for k in range(touching):
        a, b = dem[2 * k], dem[2 * k + 1]
        c.r[b] = c.r[a] + np.array([1.9 * R * math.cos(0.3 * k), 1.9 * R * math.sin(0.3 * k), 0.0])
    for k in range(floor):
        c.r[dem[2 * touching + k], 2] = 0.8 * R
    u = splitmix64(seed + 3, 3 * c.n).reshape(c.n, 3)
    v = vscale * (2.0 * u - 1.0)
    v[c.iam == SPH_BOUNDARY] = 0.0
    c.v[:] = v
    c.vhat[:] = v
    c.name = "adjoint"
    c.info.update({"vscale": vscale, "touching": touching, "floor": floor})
    return c


def adjoint_target(case: Case, seed: int = 1004, depth: float = 1e-4) -> np.ndarray:
    """Seeded target profile of the DEM heights: each DEM's initial z lowered by
    depth * U[0, 1)."""
    z = case.r[case.iam == DEM, 2]
    return z - depth * splitmix64(seed + 5, len(z))


def ensemble_cases(nbox: int = 256, seed: int = 1005, nmin: int = 4096,
                   nmax: int = 65536) -> list:
    """Independent sedimentation boxes of uneven size: per box a fluid count N
    drawn log-uniformly from [nmin, nmax] (an m x m x nz lattice with m the
    nearest integer to N^(1/3), nz = round(N / m^2)), N/16 DEM spheres, and the
    Di Felice chi-law coefficients, the contact stiffness k and the damping
    gamma each scaled by a factor drawn uniformly from [0.8, 1.2] (DEM_TIMESTEP
    recomputed from k and gamma as demConstants.h:6 does)."""
    rng = np.random.default_rng(seed)
    cases = []
    for b in range(nbox):
        N = int(math.floor(math.exp(rng.uniform(math.log(nmin), math.log(nmax)))))
        m = max(4, int(round(N ** (1.0 / 3.0))))
        nz = max(4, int(round(N / (m * m))))
        c = sediment_case(nf=m, n_dem=(m * m * nz) // 16, seed=seed + 7 * (b + 1), nz=nz)
        sd, sk, sg = rng.uniform(0.8, 1.2, size=3)
        P = c.params
        P.drag_a *= sd
        P.drag_b *= sd
        P.dem_k *= sk
        P.dem_gamma *= sg
        m_red = 0.5 * P.dem_vol * P.dem_dens
        P.dem_timestep = (1.0 / 50.0) * PI_REF / math.sqrt(
            P.dem_k / m_red - math.pow(0.5 * P.dem_gamma / m_red, 2))
        c.info.update({"box": b, "scale_drag": sd, "scale_k": sk, "scale_gamma": sg})
        cases.append(c)
    return cases


# ----------------------------------------------------------------------------
# the GPU library
# ----------------------------------------------------------------------------
class SphDemError(RuntimeError):
    """A failed C-ABI call; `code` is its SPHDEM_E_* return value."""

    def __init__(self, msg: str, code: int = 0):
        super().__init__(msg)
        self.code = code


_lib = None


def load_library(path: str = LIB_PATH):
    """Loads libsphdem.so; raises if it is missing (no CPU fallback exists)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(path):
        raise SphDemError(f"{path} is not built: run __graft_entry__.build()")
    lib = ctypes.CDLL(path)
    vp = ctypes.c_void_p
    i64 = ctypes.c_int64
    dp = ctypes.POINTER(ctypes.c_double)
    ip = ctypes.POINTER(ctypes.c_int32)
    lib.sphdem_create.argtypes = [ctypes.POINTER(Params), ctypes.c_int, i64, ctypes.POINTER(vp)]
    lib.sphdem_destroy.argtypes = [vp]
    lib.sphdem_last_error.argtypes = [vp]
    lib.sphdem_last_error.restype = ctypes.c_char_p
    lib.sphdem_output_vars.argtypes = [vp, ctypes.POINTER(Output)]
    lib.sphdem_checkpoint_size.argtypes = [vp, ctypes.POINTER(i64)]
    lib.sphdem_checkpoint_save.argtypes = [vp, ctypes.c_void_p, i64]
    lib.sphdem_checkpoint_restore.argtypes = [vp, ctypes.c_char_p, i64]
    lib.sphdem_set_params.argtypes = [vp, ctypes.POINTER(Params)]
    lib.sphdem_upload.argtypes = [vp, i64, dp, dp, dp, dp, dp, dp, ip]
    lib.sphdem_num_particles.argtypes = [vp, ctypes.POINTER(i64)]
    lib.sphdem_set_field.argtypes = [vp, ctypes.c_int, vp]
    lib.sphdem_set_field_device.argtypes = [vp, ctypes.c_int, vp]
    lib.sphdem_get_field.argtypes = [vp, ctypes.c_int, vp]
    lib.sphdem_set_globals.argtypes = [vp, ctypes.POINTER(Globals)]
    lib.sphdem_get_globals.argtypes = [vp, ctypes.POINTER(Globals)]
    for nm in ("sphdem_sync", "sphdem_step_start", "sphdem_step_middle", "sphdem_step_end",
               "sphdem_hook_dem_floor", "sphdem_hook_sum_drag", "sphdem_reset",
               "sphdem_pass_omega", "sphdem_pass_eos", "sphdem_pass_force",
               "sphdem_pass_dddt", "sphdem_pass_density", "sphdem_pass_porosity",
               "sphdem_pass_fpi", "sphdem_pass_dem_contacts"):
        getattr(lib, nm).argtypes = [vp]
    lib.sphdem_hook_body_force.argtypes = [vp, ctypes.c_uint32, dp]
    lib.sphdem_hook_freeze.argtypes = [vp, ctypes.c_uint32, ctypes.c_double]
    lib.sphdem_run_steps.argtypes = [vp, ctypes.c_int, ctypes.c_int]
    lib.sphdem_run_steps_many.argtypes = [ctypes.POINTER(vp), ctypes.c_int, ctypes.c_int,
                                          ctypes.c_int, ctypes.c_int,
                                          ctypes.POINTER(ctypes.c_int)]
    lib.sphdem_run_steps_graph.argtypes = [vp, ctypes.c_int, ctypes.c_int]
    lib.sphdem_run_steps_many_graph.argtypes = lib.sphdem_run_steps_many.argtypes
    lib.sphdem_run_steps_ensemble_graph.argtypes = lib.sphdem_run_steps_many.argtypes
    lib.sphdem_get_cell_coords.argtypes = [vp, ip]
    lib.sphdem_get_pairs.argtypes = [vp, ctypes.c_int, ctypes.POINTER(i64), ip, ip, ip, i64]
    lib.sphdem_get_pairs_device.argtypes = [vp, ctypes.c_int, ctypes.POINTER(i64), vp, vp, vp,
                                            i64]
    lib.sphdem_reduce_dt.argtypes = [vp, dp]
    lib.sphdem_num_deleted.argtypes = [vp, ctypes.POINTER(i64)]
    lib.sphdem_set_tile_shape.argtypes = [vp, ctypes.c_int, ctypes.c_int]
    lib.sphdem_get_tile_shape.argtypes = [vp, ctypes.POINTER(ctypes.c_int),
                                          ctypes.POINTER(ctypes.c_int)]
    lib.sphdem_tune_tile_shape.argtypes = [vp, ctypes.c_int, ctypes.c_int,
                                           ctypes.POINTER(ctypes.c_int),
                                           ctypes.POINTER(ctypes.c_int), dp]
    lib.sphdem_timing_begin.argtypes = [vp, ctypes.c_char_p]
    lib.sphdem_timing_end.argtypes = [vp, dp, ctypes.POINTER(i64)]
    _lib = lib
    return lib


def run_many(sims: list, nsteps: int, hookset: int | None = None, workers: int = 0,
             graph: bool | str = False):
    """Advances independent boxes together (sphdem_run_steps_many; with
    graph=True sphdem_run_steps_many_graph, each box replaying its own graph;
    with graph="ensemble" sphdem_run_steps_ensemble_graph, one graph over all
    boxes): each Sim by nsteps steps, `workers` host threads stepping their
    share of the boxes."""
    if graph not in (False, True, "ensemble"):
        raise ValueError(f"graph must be False, True or 'ensemble', not {graph!r}")
    if not sims:
        return
    lib = load_library()
    if hookset is None:
        hookset = sims[0].case.hookset
        if any(s.case.hookset != hookset for s in sims):
            raise ValueError("boxes with different hook sets need an explicit hookset")
    arr = (ctypes.c_void_p * len(sims))(*[s.ctx.value for s in sims])
    failed = ctypes.c_int(-1)
    fn = (lib.sphdem_run_steps_ensemble_graph if graph == "ensemble" else
          lib.sphdem_run_steps_many_graph if graph else lib.sphdem_run_steps_many)
    rc = fn(arr, len(sims), nsteps, hookset, workers, ctypes.byref(failed))
    if rc != 0:
        msg = b""
        if failed.value >= 0:
            msg = lib.sphdem_last_error(sims[failed.value].ctx) or b""
        raise SphDemError(f"{fn.__name__} failed ({rc}) in box {failed.value}: "
                          f"{msg.decode(errors='replace')}")


def _dptr(a):
    return None if a is None else a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


class Sim:
    """One device context holding one case."""

    def __init__(self, case: Case, device: int = 0, precision: int = FP64, capacity: int = 0):
        self.lib = load_library()
        self.case = case
        P = case.params.copy()
        P.precision = precision
        self.params = P
        h = ctypes.c_void_p()
        self._check(self.lib.sphdem_create(ctypes.byref(P), device,
                                           capacity or case.n, ctypes.byref(h)), None)
        self.ctx = h
        r = np.ascontiguousarray(case.r, dtype=np.float64)
        v = np.ascontiguousarray(case.v, dtype=np.float64)
        vh = np.ascontiguousarray(case.vhat, dtype=np.float64)
        m = np.ascontiguousarray(case.mass, dtype=np.float64)
        hh = np.ascontiguousarray(case.h, dtype=np.float64)
        d = np.ascontiguousarray(case.dens, dtype=np.float64)
        iam = np.ascontiguousarray(case.iam, dtype=np.int32)
        self._check(self.lib.sphdem_upload(self.ctx, case.n, _dptr(r), _dptr(v), _dptr(vh),
                                           _dptr(m), _dptr(hh), _dptr(d),
                                           iam.ctypes.data_as(ctypes.POINTER(ctypes.c_int32))))
        g = Globals()
        g.maxdt = P.maxdt
        self.set_globals(g)

    def _check(self, rc, ctx="self"):
        if rc != 0:
            msg = b""
            if ctx is not None and getattr(self, "ctx", None):
                msg = self.lib.sphdem_last_error(self.ctx) or b""
            raise SphDemError(f"sphdem call failed ({rc}): {msg.decode(errors='replace')}", rc)

    def close(self):
        if getattr(self, "ctx", None):
            self.lib.sphdem_destroy(self.ctx)
            self.ctx = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # ---- state
    def get(self, name: str) -> np.ndarray:
        n = self.case.n
        if name in INT_FIELDS:
            out = np.empty(n, dtype=np.int32)
        elif name in VECTOR_FIELDS:
            out = np.empty((n, 3), dtype=np.float64)
        else:
            out = np.empty(n, dtype=np.float64)
        self._check(self.lib.sphdem_get_field(self.ctx, FIELD[name], out.ctypes.data))
        return out

    def set(self, name: str, value: np.ndarray):
        dt = np.int32 if name in INT_FIELDS else np.float64
        a = np.ascontiguousarray(value, dtype=dt)
        self._check(self.lib.sphdem_set_field(self.ctx, FIELD[name], a.ctypes.data))

    def set_device(self, name: str, ptr: int):
        """sphdem_set_field_device: `ptr` addresses n (n*3 for a vector field)
        float64 values in original order in this context's device memory, with
        the writes that made them complete."""
        self._check(self.lib.sphdem_set_field_device(self.ctx, FIELD[name], ctypes.c_void_p(ptr)))

    def globals(self) -> Globals:
        g = Globals()
        self._check(self.lib.sphdem_get_globals(self.ctx, ctypes.byref(g)))
        return g

    def set_globals(self, g: Globals):
        self._check(self.lib.sphdem_set_globals(self.ctx, ctypes.byref(g)))

    def set_params(self, P: Params):
        self.params = P.copy()
        self._check(self.lib.sphdem_set_params(self.ctx, ctypes.byref(self.params)))

    def sync(self):
        self._check(self.lib.sphdem_sync(self.ctx))

    # ---- loop
    def start(self):
        self._check(self.lib.sphdem_step_start(self.ctx))

    def middle(self):
        self._check(self.lib.sphdem_step_middle(self.ctx))

    def end(self):
        self._check(self.lib.sphdem_step_end(self.ctx))

    def run(self, nsteps: int, hookset: int | None = None):
        hs = self.case.hookset if hookset is None else hookset
        self._check(self.lib.sphdem_run_steps(self.ctx, nsteps, hs))

    def run_graph(self, nsteps: int, hookset: int | None = None):
        """sphdem_run_steps_graph: the same steps replayed from a captured graph."""
        hs = self.case.hookset if hookset is None else hookset
        self._check(self.lib.sphdem_run_steps_graph(self.ctx, nsteps, hs))

    def hook_body_force(self, type_mask: int, accel):
        a = np.ascontiguousarray(accel, dtype=np.float64)
        self._check(self.lib.sphdem_hook_body_force(self.ctx, type_mask, _dptr(a)))

    def hook_dem_floor(self):
        self._check(self.lib.sphdem_hook_dem_floor(self.ctx))

    def hook_freeze(self, type_mask: int, t: float):
        self._check(self.lib.sphdem_hook_freeze(self.ctx, type_mask, t))

    def hook_sum_drag(self):
        self._check(self.lib.sphdem_hook_sum_drag(self.ctx))

    # ---- parity
    def reset(self):
        self._check(self.lib.sphdem_reset(self.ctx))

    def cells(self) -> np.ndarray:
        out = np.empty((self.case.n, 3), dtype=np.int32)
        self._check(self.lib.sphdem_get_cell_coords(
            self.ctx, out.ctypes.data_as(ctypes.POINTER(ctypes.c_int32))))
        return out

    def output_vars(self) -> dict:
        """calcOutputVars diagnostics (sphdem_output_vars) as a dict."""
        o = Output()
        self._check(self.lib.sphdem_output_vars(self.ctx, ctypes.byref(o)))
        d = {k: getattr(o, k) for k, _ in Output._fields_}
        d["linMom"] = list(o.linMom)
        return d

    def checkpoint(self) -> bytes:
        """Full-state checkpoint (sphdem_checkpoint_save): continuing from it
        with restore() reproduces the run bit for bit."""
        size = ctypes.c_int64()
        self._check(self.lib.sphdem_checkpoint_size(self.ctx, ctypes.byref(size)))
        buf = (ctypes.c_char * size.value)()
        self._check(self.lib.sphdem_checkpoint_save(self.ctx, buf, size.value))
        return bytes(buf)

    def restore(self, blob: bytes):
        self._check(self.lib.sphdem_checkpoint_restore(self.ctx, blob, len(blob)))

    def pair_count(self, kind: int) -> int:
        """Number of pairs of one neighbour search (the cached lists for SPH)."""
        cnt = ctypes.c_int64()
        self._check(self.lib.sphdem_get_pairs(self.ctx, kind, ctypes.byref(cnt), None, None,
                                              None, 0))
        return cnt.value

    def pairs(self, kind: int) -> np.ndarray:
        cnt = ctypes.c_int64()
        self._check(self.lib.sphdem_get_pairs(self.ctx, kind, ctypes.byref(cnt), None, None,
                                              None, 0))
        n = cnt.value
        a = np.empty(max(n, 1), dtype=np.int32)
        b = np.empty(max(n, 1), dtype=np.int32)
        im = np.empty(max(n, 1), dtype=np.int32)
        ip = ctypes.POINTER(ctypes.c_int32)
        self._check(self.lib.sphdem_get_pairs(self.ctx, kind, ctypes.byref(cnt),
                                              a.ctypes.data_as(ip), b.ctypes.data_as(ip),
                                              im.ctypes.data_as(ip), n))
        return np.stack([a[:n], b[:n], im[:n]], axis=1)

    def pass_(self, name: str):
        self._check(getattr(self.lib, "sphdem_pass_" + name)(self.ctx))

    def reduce_dt(self) -> float:
        v = ctypes.c_double()
        self._check(self.lib.sphdem_reduce_dt(self.ctx, ctypes.byref(v)))
        return v.value

    def num_deleted(self) -> int:
        v = ctypes.c_int64()
        self._check(self.lib.sphdem_num_deleted(self.ctx, ctypes.byref(v)))
        return v.value

    # ---- tile shape (sphdem_set/get/tune_tile_shape)
    def set_tile_shape(self, seg_particles: int, walker_block: int):
        self._check(self.lib.sphdem_set_tile_shape(self.ctx, seg_particles, walker_block))

    def tile_shape(self) -> tuple[int, int]:
        s, b = ctypes.c_int(), ctypes.c_int()
        self._check(self.lib.sphdem_get_tile_shape(self.ctx, ctypes.byref(s), ctypes.byref(b)))
        return s.value, b.value

    def tune_tile_shape(self, steps: int = 2, hookset: int | None = None):
        """Times each candidate shape over `steps` steps from the current state
        (restored afterwards), keeps the fastest; returns (segment particles,
        walker block, ms per step)."""
        hs = self.case.hookset if hookset is None else hookset
        s, b, ms = ctypes.c_int(), ctypes.c_int(), ctypes.c_double()
        self._check(self.lib.sphdem_tune_tile_shape(self.ctx, steps, hs, ctypes.byref(s),
                                                    ctypes.byref(b), ctypes.byref(ms)))
        return s.value, b.value, ms.value

    # ---- measurement: HIP events around each launch of one kernel
    def timing_begin(self, kernel: str):
        self._check(self.lib.sphdem_timing_begin(self.ctx, kernel.encode()))

    def timing_end(self) -> tuple[float, int]:
        ms = ctypes.c_double()
        n = ctypes.c_int64()
        self._check(self.lib.sphdem_timing_end(self.ctx, ctypes.byref(ms), ctypes.byref(n)))
        return ms.value, n.value
