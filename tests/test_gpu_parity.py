"""Parity of the gfx950 library (through its C ABI) with the CPU oracle on
identical seeded inputs.

Bar (FP64 form): cell coordinates and neighbour-pair sets bit-exact; every
field of a pass within 1e-12 normwise relative error
(max_i |gpu_i - oracle_i| / max_i |oracle_i| over the particles the pass
writes); multi-step runs within the tolerances stated per test."""
import numpy as np
import pytest

import oracle
import sphdem as sd

pytestmark = pytest.mark.gpu

TOL_PASS = 1e-12


def normwise(a, b):
    a = np.asarray(a, float)
    b = np.asarray(b, float)
    scale = np.abs(b).max() if b.size else 0.0
    if scale == 0.0:
        return float(np.abs(a).max()) if a.size else 0.0
    return float(np.abs(a - b).max() / scale)


def as_set(p):
    return set(map(tuple, np.asarray(p, dtype=np.int64).tolist()))


def assert_same_pairs(g, o):
    gs, os_ = as_set(g), as_set(o)
    extra, missing = gs - os_, os_ - gs
    assert not extra and not missing, (
        f"{len(extra)} pairs only on the GPU (e.g. {sorted(extra)[:5]}), "
        f"{len(missing)} only in the oracle (e.g. {sorted(missing)[:5]})")
    assert len(g) == len(o)


def dam_case():
    return sd.dambreak_case(nf=12)


def sed_case():
    return sd.sediment_case(nf=12, n_dem=60)


def contact_case():
    """sediment with DEM spheres pushed into overlapping pairs."""
    c = sd.sediment_case(nf=12, n_dem=40, seed=77)
    dem = np.where(c.iam == sd.DEM)[0]
    R = c.params.dem_radius
    r = c.r.copy()
    rng = np.random.default_rng(2)
    for k in range(0, len(dem) - 1, 2):
        d = rng.normal(size=3)
        d /= np.linalg.norm(d)
        r[dem[k + 1]] = r[dem[k]] + 1.9 * R * d
    c.r = r
    return c


def closedbox_case():
    """The large mixed-precision workload's geometry at 12^3: the dam-break box
    closed on every side with the SPH-DEM flag set and DEM spheres in the upper
    half of the fluid block (1.6 % of the fluid count, as 65,536 of 160^3)."""
    return sd.dambreak_case(nf=12, with_dem=28, liq_dem_flags=True)


CASES = {"dambreak": dam_case, "sediment": sed_case, "closedbox": closedbox_case}


@pytest.fixture(scope="module", params=list(CASES))
def case(request):
    return CASES[request.param]()


def both(case):
    return sd.Sim(case), oracle.Oracle(case)


def perturb(case, sim, orc, seed=11):
    """Identical nontrivial state in both: velocities, densities, previous DEM drag."""
    rng = np.random.default_rng(seed)
    n = case.n
    vh = rng.normal(size=(n, 3)) * 0.05
    vh[case.iam == sd.SPH_BOUNDARY] = 0.0
    dens = case.dens * (1 + 0.002 * rng.normal(size=n))
    dens[case.iam == sd.DEM] = case.dens[case.iam == sd.DEM]
    fd = np.zeros((n, 3))
    fd[case.iam == sd.DEM] = rng.normal(size=((case.iam == sd.DEM).sum(), 3))
    for x in (sim, orc):
        x.set("vhat", vh)
        x.set("v", vh)
        x.set("dens", dens)
        x.set("fdrag", fd)
    orc.refresh_ghosts()


def fresh_newdt(sim, orc, value):
    for x in (sim, orc):
        g = x.globals()
        g.newDt = value
        x.set_globals(g)


def check_fields(sim, orc, names, sel, tol=TOL_PASS):
    errs = {}
    for nm in names:
        errs[nm] = normwise(sim.get(nm)[sel], orc.get(nm)[sel])
    bad = {k: v for k, v in errs.items() if not v <= tol}
    assert not bad, f"fields above {tol}: {bad} (all: {errs})"
    return errs


# ------------------------------------------------------------ neighbour search
def test_cell_coordinates_bit_exact(case):
    sim, orc = both(case)
    assert np.array_equal(sim.cells(), orc.cells())


@pytest.mark.parametrize("kind", [sd.PAIRS_SPH, sd.PAIRS_DEM, sd.PAIRS_SPH2DEM, sd.PAIRS_DEM2SPH])
def test_pair_sets_bit_exact(case, kind):
    if kind != sd.PAIRS_SPH and not (case.params.flags & sd.F_LIQ_DEM):
        pytest.skip("no DEM in this case")
    sim, orc = both(case)
    assert_same_pairs(sim.pairs(kind), orc.pairs(kind))


def test_contact_pairs_bit_exact():
    c = contact_case()
    sim, orc = both(c)
    g, o = sim.pairs(sd.PAIRS_DEM), orc.pairs(sd.PAIRS_DEM)
    assert len(o) > 0
    assert_same_pairs(g, o)


def lists_by_particle(p):
    """{a: [(b, image), ...]} in the order the pair array holds them."""
    out = {}
    for a, b, im in np.asarray(p, dtype=np.int64).tolist():
        out.setdefault(a, []).append((b, im))
    return out


def test_sph_list_order(case):
    """The cached SPH lists hold the neighbours in the reference's order: cells
    x-outer, then y, then z; in a cell the real particles in container order,
    then the periodic ghosts in the order they are received (receive slot,
    then the origin's container order)."""
    sim, orc = both(case)
    g = lists_by_particle(sim.pairs(sd.PAIRS_SPH))
    o = lists_by_particle(orc.pairs(sd.PAIRS_SPH))
    assert g.keys() == o.keys()
    with_images = 0
    for a, lo in o.items():
        assert g[a] == lo, f"particle {a}: list order differs"
        with_images += any(im != 13 for _, im in lo)
    if case.params.periodic[0]:
        assert with_images > 0


# ------------------------------------------------------------ per-pass parity
def test_middle_passes(case):
    """middle()'s passes on identical state: coupling (porosity, Shepard sum,
    DEM->SPH drag), omega, EOS, force and the pair dt conditions."""
    sim, orc = both(case)
    perturb(case, sim, orc)
    fresh_newdt(sim, orc, case.params.maxdt)
    liq = bool(case.params.flags & sd.F_LIQ_DEM)
    sph = (case.iam == sd.SPH) | (case.iam == sd.SPH_BOUNDARY)
    dem = case.iam == sd.DEM
    if liq:
        sim.pass_("porosity")
        orc.pass_("porosity")
        check_fields(sim, orc, ["shepSum"], dem)
        check_fields(sim, orc, ["porosity", "fdrag", "f"], sph)
    if case.params.flags & sd.F_VAR_H_CORRECTION:
        sim.pass_("omega")
        orc.pass_("omega")
        check_fields(sim, orc, ["omega"], sph)
    sim.pass_("eos")
    orc.pass_("eos")
    check_fields(sim, orc, ["press", "pdr2"], sph)
    sim.pass_("force")
    orc.pass_("force")
    check_fields(sim, orc, ["f", "fp", "fv"], sph)
    assert sim.reduce_dt() == pytest.approx(orc.new_dt(), rel=1e-15)


def test_end_passes(case):
    """end()'s passes: omega, dddt, then FPI with the Di Felice drag."""
    sim, orc = both(case)
    perturb(case, sim, orc, seed=12)
    sph = (case.iam == sd.SPH) | (case.iam == sd.SPH_BOUNDARY)
    dem = case.iam == sd.DEM
    liq = bool(case.params.flags & sd.F_LIQ_DEM)
    if liq:  # forces and porosity the SPH side hands to the DEM gather
        rng = np.random.default_rng(4)
        for nm in ("f", "fp", "fv", "fb"):
            x = rng.normal(size=(case.n, 3))
            sim.set(nm, x)
            orc.set(nm, x)
        por = np.clip(1 - 0.2 * rng.random(case.n), 0.37, 1)
        sim.set("porosity", por)
        orc.set("porosity", por)
        orc.refresh_ghosts()
    if case.params.flags & sd.F_VAR_H_CORRECTION:
        sim.pass_("omega")
        orc.pass_("omega")
    sim.pass_("dddt")
    orc.pass_("dddt")
    check_fields(sim, orc, ["dddt"], sph)
    if liq:
        sim.pass_("fpi")
        orc.pass_("fpi")
        check_fields(sim, orc, ["fdrag", "f", "fvel", "porosity", "shepSum", "fb", "fv"], dem)


def test_dem_contacts_pass():
    c = contact_case()
    sim, orc = both(c)
    rng = np.random.default_rng(9)
    vh = rng.normal(size=(c.n, 3)) * 0.01
    for x in (sim, orc):
        x.set("vhat", vh)
    orc.refresh_ghosts()
    dem = c.iam == sd.DEM
    sim.pass_("dem_contacts")
    orc.pass_("dem_contacts")
    errs = check_fields(sim, orc, ["fp", "f"], dem)
    assert np.abs(orc.get("fp")[dem]).max() > 0


def test_density_reinit_pass():
    c = dam_case()
    sim, orc = both(c)
    perturb(c, sim, orc, seed=13)
    sim.pass_("density")
    orc.pass_("density")
    sph = c.iam != sd.DEM
    check_fields(sim, orc, ["h", "dens"], sph)


# ------------------------------------------------------------ whole steps
@pytest.mark.parametrize("nsteps", [1, 3])
def test_steps(case, nsteps):
    """run.cpp's loop (sphStep++, hooks, start/middle/end) for a few steps."""
    sim, orc = both(case)
    sim.run(nsteps)
    orc.run(nsteps)
    gs, go = sim.globals(), orc.globals()
    assert gs.dt == pytest.approx(go.dt, rel=1e-12)
    assert gs.time == pytest.approx(go.time, rel=1e-12)
    alive = orc.get("iam") != -1
    tol = 1e-11 if nsteps == 1 else 1e-9
    names = ["r", "v", "vhat", "dens", "h", "f", "press"]
    if case.params.flags & sd.F_LIQ_DEM:
        names += ["porosity", "fdrag", "shepSum"]
    check_fields(sim, orc, names, alive, tol)
    # end()'s dddt walk and, under VAR_H, the omega summed in it
    iam = orc.get("iam")
    sphg = (iam == sd.SPH) | (iam == sd.SPH_BOUNDARY)
    extra = ["dddt"] + (["omega"] if case.params.flags & sd.F_VAR_H_CORRECTION else [])
    check_fields(sim, orc, extra, sphg, tol)
    # the neighbour sets of the evolved state (lists rebuilt by a reset: the
    # cached ones predate end()'s driftRest, which changes h) are identical
    sim.reset()
    orc.reset()
    assert_same_pairs(sim.pairs(sd.PAIRS_SPH), orc.pairs(sd.PAIRS_SPH))


VISC_VARIANTS = {"cleary": sd.F_VISC_CLEARY, "monaghan": sd.F_VISC_MONAGHAN,
                 "monaghan+artificial": sd.F_VISC_MONAGHAN | sd.F_VISC_ARTIFICIAL}


def visc_case(flags):
    c = dam_case()
    c.params.flags = (c.params.flags & ~(sd.F_VISC_MORRIS | sd.F_VISC_ARTIFICIAL)) | flags
    return c


@pytest.mark.parametrize("variant", sorted(VISC_VARIANTS))
def test_viscosity_variants(variant):
    """calcViscForce's Cleary and Monaghan branches (sph.h:598-630) and their dt
    conditions: the force pass on identical state, then three whole steps."""
    c = visc_case(VISC_VARIANTS[variant])
    print(f"viscosity variant {variant}", flush=True)
    sim, orc = both(c)
    perturb(c, sim, orc, seed=31)
    fresh_newdt(sim, orc, c.params.maxdt)
    sph = (c.iam == sd.SPH) | (c.iam == sd.SPH_BOUNDARY)
    for x in (sim, orc):
        x.pass_("eos")
        x.pass_("force")
    check_fields(sim, orc, ["f", "fp", "fv"], sph)
    assert np.abs(orc.get("fv")[sph]).max() > 0
    assert sim.reduce_dt() == pytest.approx(orc.new_dt(), rel=1e-15)
    sim, orc = both(c)
    sim.run(3)
    orc.run(3)
    assert sim.globals().dt == pytest.approx(orc.globals().dt, rel=1e-12)
    alive = orc.get("iam") != -1
    check_fields(sim, orc, ["r", "v", "vhat", "dens", "h", "f", "press"], alive, 1e-9)


def _with_flags(c, add=0, damptime=None):
    c.params.flags |= add
    if damptime is not None:
        c.params.damptime = damptime
    return c


def _without_custom_wall(c):
    """LIQ_DEM_CUSTOM_WALL_CONTACTS off: the reference then runs
    calcDEMWallContacts, whose scan of the DEM list finds no wall particle."""
    c.params.flags &= ~sd.F_CUSTOM_WALL
    return c


def floor_case():
    """sediment with DEM spheres overlapping the z = RMIN floor (centre at
    0.8 R): the sediment3D floor contact (addCustomBoundaries,
    sediment3D/customSim.cpp:67-81) acts on them from the first step."""
    c = sd.sediment_case(nf=12, n_dem=40, seed=78)
    dem = np.where(c.iam == sd.DEM)[0]
    R = c.params.dem_radius
    c.r[dem[:12], 2] = 0.8 * R
    return c


SWITCH_CASES = {
    "dem_floor": floor_case,
    "dem_block": lambda: _with_flags(contact_case(), sd.F_DEM_BLOCK),
    "no_custom_wall": lambda: _without_custom_wall(floor_case()),
    "const_h": lambda: _with_flags(dam_case(), sd.F_CONST_H),
    "damptime": lambda: _with_flags(dam_case(), damptime=1),
    "fixed_dem": lambda: _with_flags(contact_case(), sd.F_FIXED_DEM),
    "no_dem_contacts": lambda: _with_flags(contact_case(), sd.F_NO_DEM_CONTACTS),
}


@pytest.mark.parametrize("name", sorted(SWITCH_CASES))
def test_step_switches(name):
    """CONST_H (h kept), DAMPTIME (velocities x 0.98 while time < DAMPTIME,
    sphIncompress.h:716-721), FIXED_DEM (DEM velocities restored after every
    kick, no contacts), NO_DEM_CONTACTS, DEM_BLOCK (every DEM force the mean,
    sphIncompress.cpp:231-238), spheres on the custom floor, and the flag set
    without LIQ_DEM_CUSTOM_WALL_CONTACTS: three whole steps against the
    oracle."""
    c = SWITCH_CASES[name]()
    print(f"switch case {name}", flush=True)
    sim, orc = both(c)
    sim.run(3)
    orc.run(3)
    assert sim.globals().dt == pytest.approx(orc.globals().dt, rel=1e-12)
    assert sim.globals().time == pytest.approx(orc.globals().time, rel=1e-12)
    alive = orc.get("iam") != -1
    names = ["r", "v", "vhat", "dens", "h", "f", "press"]
    if c.params.flags & sd.F_LIQ_DEM:
        names += ["porosity", "fdrag", "fp"]
    check_fields(sim, orc, names, alive, 1e-9)


def test_reinit_every_step():
    """REINIT_DENS_EVERY_N_STEPS = 1 (as damBreak sets): 10 density passes per step."""
    c = dam_case()
    c.params.reinit_every = 1
    sim, orc = both(c)
    sim.run(2)
    orc.run(2)
    alive = orc.get("iam") != -1
    check_fields(sim, orc, ["dens", "h", "r", "f"], alive, 1e-10)


def test_out_of_range_deletion_matches():
    c = sd.dambreak_case(nf=8)
    sim, orc = both(c)
    assert sim.num_deleted() == orc.num_deleted() > 0
    alive = orc.get("iam") != -1
    assert np.array_equal(sim.get("iam") != sd.DELETED, alive)
    sim.run(1)
    orc.run(1)
    check_fields(sim, orc, ["r", "dens", "f"], alive, 1e-11)


def test_periodic_wrap_matches():
    c = sd.sediment_case(nf=12, n_dem=10)
    sim, orc = both(c)
    r = c.r.copy()
    r[0, 0] = 1.0 + 0.3 / 12
    r[1, 1] = -0.2 / 12
    for x in (sim, orc):
        x.set("r", r)
        x.reset()
    assert np.array_equal(sim.get("r"), orc.get("r"))
    assert np.array_equal(sim.cells(), orc.cells())


def test_larger_sediment_steps():
    """16^3 fluid, 200 DEM: more cells per column, ghost columns on all four
    periodic faces and several segments per column; two whole steps, then the
    pair set and the list order of the evolved state."""
    c = sd.sediment_case(nf=16, n_dem=200, seed=5)
    sim, orc = both(c)
    sim.run(2)
    orc.run(2)
    alive = orc.get("iam") != -1
    check_fields(sim, orc, ["r", "v", "vhat", "dens", "h", "f", "porosity", "fdrag", "shepSum"],
                 alive, 1e-10)
    sim.reset()
    orc.reset()
    g = lists_by_particle(sim.pairs(sd.PAIRS_SPH))
    o = lists_by_particle(orc.pairs(sd.PAIRS_SPH))
    assert g == o
