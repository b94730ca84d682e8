"""Pins the CPU oracle (the checker of the GPU library) against what the
mathematics fixes: closed forms, brute-force neighbour searches, conservation
laws and hand-derived trajectories.  The reference itself cannot be compiled
in this image (vect.h:7 needs the absent GSL headers), so these checks are
what the oracle rests on.  CPU only."""
import math

import numpy as np
import pytest

import oracle
import sphdem as sd


# ----------------------------------------------------------------- helpers
def brute_pairs(r, h, iam, sel_a, sel_b, P, radius_fn, strict=False):
    """All (a, b, image) with the reference's arithmetic for image positions:
    img = (RMAX-RMIN) + r_b or (RMIN-RMAX) + r_b (dataLL.cpp:469-484),
    d = r_a - img, len2 = d0*d0 + (d1*d1 + d2*d2)."""
    A = np.where(sel_a)[0]
    B = np.where(sel_b)[0]
    shifts = [[0] + ([-1, 1] if P.periodic[i] else []) for i in range(3)]
    out = []
    for sx in shifts[0]:
        for sy in shifts[1]:
            for sz in shifts[2]:
                s = (sx, sy, sz)
                img = r[B].copy()
                for i in range(3):
                    if s[i] == 1:
                        img[:, i] = (P.rmax[i] - P.rmin[i]) + img[:, i]
                    elif s[i] == -1:
                        img[:, i] = (P.rmin[i] - P.rmax[i]) + img[:, i]
                code = (sx + 1) + 3 * (sy + 1) + 9 * (sz + 1)
                for c0 in range(0, len(A), 512):
                    a = A[c0:c0 + 512]
                    d = r[a][:, None, :] - img[None, :, :]
                    l2 = d[..., 0] * d[..., 0] + (d[..., 1] * d[..., 1] + d[..., 2] * d[..., 2])
                    lim = radius_fn(a, B)
                    m = (l2 < lim) if strict else (l2 <= lim)
                    if code == 13:
                        m &= a[:, None] != B[None, :]
                    ia, ib = np.nonzero(m)
                    out.append(np.stack([a[ia], B[ib], np.full(len(ia), code)], axis=1))
    return np.concatenate(out) if out else np.zeros((0, 3), int)


def as_set(p):
    return set(map(tuple, np.asarray(p, dtype=np.int64).tolist()))


@pytest.fixture(scope="module")
def dam():
    return sd.dambreak_case(nf=12)


@pytest.fixture(scope="module")
def sed():
    return sd.sediment_case(nf=12, n_dem=60)


# ------------------------------------------------------- kernel family
def test_wendland_normalisation_and_support(dam):
    o = oracle.Oracle(dam)
    h = 0.7
    q = np.linspace(0.0, 2.0, 20001)
    w = np.array([o.W(x, h) for x in q])
    integral = np.trapezoid(4 * math.pi * (q * h) ** 2 * w, q * h)
    # WCON uses PI = 3.14159265 (customConstants.h:11), so the integral is
    # pi/3.14159265 = 1 + 1.1e-9
    assert abs(integral - math.pi / sd.PI_REF) < 1e-8
    assert o.W(2.0, h) == 0.0 and o.W(2.5, h) == 0.0 and o.F(2.5, h) == 0.0
    assert o.W(0.0, h) == pytest.approx(21.0 / (16.0 * sd.PI_REF) / h ** 3, rel=1e-15)


def test_F_is_derivative_of_K(dam):
    o = oracle.Oracle(dam)
    for h in (0.3, 1.7):
        for q in (0.1, 0.5, 1.0, 1.5, 1.9):
            eps = 1e-6
            dKdq = (o.K(q + eps, h) - o.K(q - eps, h)) / (2 * eps)
            # gradW = dr F with F = dK/dq / (q h^(NDIM+2))
            assert o.F(q, h) == pytest.approx(dKdq / (q * h ** 5), rel=1e-7)


def test_pdWh_is_h_derivative_of_W(dam):
    o = oracle.Oracle(dam)
    for h in (0.4, 1.3):
        for r in (0.0, 0.1, 0.5, 0.7):
            eps = 1e-7
            num = (o.W(r / (h + eps), h + eps) - o.W(r / (h - eps), h - eps)) / (2 * eps)
            assert o.pdWh(r / h, h) == pytest.approx(num, rel=1e-6, abs=1e-9)


# ------------------------------------------------------- neighbour searches
def test_sph_pairs_match_brute_force(dam):
    o = oracle.Oracle(dam)
    got = o.pairs(sd.PAIRS_SPH)
    r, h, iam = dam.r, dam.h, dam.iam
    sph = iam != sd.DEM
    lim = lambda a, B: np.maximum(h[a][:, None], h[B][None, :]) * 2.0
    exp = brute_pairs(r, h, iam, sph, sph, dam.params, lambda a, B: lim(a, B) ** 2)
    assert len(got) == len(exp)
    assert as_set(got) == as_set(exp)
    # ordering: lists are per particle in x-outer cell order, insertion order in a cell
    assert np.all(np.diff(got[:, 0]) >= 0)


def test_periodic_pairs_match_brute_force(sed):
    o = oracle.Oracle(sed)
    assert o.num_ghosts() > 0
    got = o.pairs(sd.PAIRS_SPH)
    r, h, iam = sed.r, sed.h, sed.iam
    sph = iam != sd.DEM
    lim = lambda a, B: (2.0 * np.maximum(h[a][:, None], h[B][None, :])) ** 2
    exp = brute_pairs(r, h, iam, sph, sph, sed.params, lim)
    assert as_set(got) == as_set(exp)
    assert np.any(got[:, 2] != 13)  # images across the periodic faces occur


def test_dem_dem_pairs_match_brute_force():
    c = sd.sediment_case(nf=12, n_dem=400, seed=1234)
    o = oracle.Oracle(c)
    got = o.pairs(sd.PAIRS_DEM)
    dem = c.iam == sd.DEM
    lim = lambda a, B: (2.0 * np.maximum(c.h[a][:, None], c.h[B][None, :])) ** 2
    exp = brute_pairs(c.r, c.h, c.iam, dem, dem, c.params, lim)
    assert as_set(got) == as_set(exp)


def coupling_rule_brute(case, o):
    """calcNeighboursCoupling's SPH-side rule (dataLL.cpp:1154-1190) evaluated
    over all (SPH, DEM image) pairs instead of by a cell scan."""
    P = case.params
    r = case.r
    Rc, DSR = P.coupling_radius, P.dem_search_radius
    nc = int(math.ceil(0.5 * Rc / DSR))
    thr = (0.5 * Rc / DSR) ** 2
    e = 2 * DSR
    s = 2 * Rc / DSR

    def cell(x):
        return ((x - np.array(P.proc_domain[0::2])) / e + s - 1).astype(np.int64)

    A = np.where(case.iam != sd.DEM)[0]
    B = np.where(case.iam == sd.DEM)[0]
    ca = cell(r[A])
    out = set()
    shifts = [[0] + ([-1, 1] if P.periodic[i] else []) for i in range(3)]
    for sx in shifts[0]:
        for sy in shifts[1]:
            for sz in shifts[2]:
                sh = (sx, sy, sz)
                img = r[B].copy()
                for i in range(3):
                    if sh[i] == 1:
                        img[:, i] = (P.rmax[i] - P.rmin[i]) + img[:, i]
                    elif sh[i] == -1:
                        img[:, i] = (P.rmin[i] - P.rmax[i]) + img[:, i]
                cb = cell(img)
                code = (sx + 1) + 3 * (sy + 1) + 9 * (sz + 1)
                dc = cb[None, :, :] - ca[:, None, :]
                inr = np.all(np.abs(dc) <= nc, axis=2)
                plus = np.sum((dc + 1) ** 2, axis=2)
                minus = np.sum((dc - 1) ** 2, axis=2)
                d = r[A][:, None, :] - img[None, :, :]
                l2 = d[..., 0] * d[..., 0] + (d[..., 1] * d[..., 1] + d[..., 2] * d[..., 2])
                take = inr & ((plus < thr) | ((minus < thr) & (l2 < Rc ** 2)))
                ia, ib = np.nonzero(take)
                out |= set(zip(A[ia].tolist(), B[ib].tolist(), [code] * len(ia)))
    return out


def test_coupling_pairs_match_cell_rule(sed):
    o = oracle.Oracle(sed)
    got = o.pairs(sd.PAIRS_SPH2DEM)
    exp = coupling_rule_brute(sed, o)
    # the reference only has DEM images within LIQ_DEM_COUPLING_RADIUS of a
    # periodic face; whole-cell ("addAll") members beyond that band exist in
    # the rule but not as ghosts
    got_s = as_set(got)
    assert got_s <= exp
    missing = exp - got_s
    r = sed.r
    P = sed.params
    for (a, b, code) in missing:
        assert code != 13
        d = r[a] - r[b]
        assert np.sqrt(d @ d) > 0  # image pair: ghost never created
    # every pair within the coupling radius is present
    Rc = P.coupling_radius
    for (a, b, code) in exp - got_s:
        sh = [(code % 3) - 1, (code // 3) % 3 - 1, code // 9 - 1]
        img = r[b] + np.array([sh[i] * (P.rmax[i] - P.rmin[i]) for i in range(3)])
        assert np.linalg.norm(r[a] - img) >= Rc


def test_dem_side_coupling_matches_brute_force(sed):
    o = oracle.Oracle(sed)
    got = o.pairs(sd.PAIRS_DEM2SPH)
    h = sed.h
    dem = sed.iam == sd.DEM
    sph = ~dem
    lim = lambda a, B: (2.0 * np.maximum(h[a][:, None], h[B][None, :])) ** 2
    exp = brute_pairs(sed.r, h, sed.iam, dem, sph, sed.params, lim)
    assert as_set(got) == as_set(exp)


def test_cell_coordinates_formula(dam, sed):
    for c in (dam, sed):
        o = oracle.Oracle(c)
        got = o.cells()
        P = c.params
        cmin = np.array(P.proc_domain[0::2])
        ssr = o.search_radius()
        sph = c.iam != sd.DEM
        exp = ((c.r[sph] - cmin) / (2.0 * ssr) + 3).astype(np.int32)
        assert np.array_equal(got[sph], exp)
        if np.any(~sph):
            e = 2 * P.dem_search_radius
            s = 2 * P.coupling_radius / P.dem_search_radius
            expd = (((c.r[~sph] - cmin) / e + s) - 1).astype(np.int32)
            assert np.array_equal(got[~sph], expd)


def test_out_of_range_particles_are_deleted():
    # nf = 8: 0.25 L < 3 dx, so the outer wall layers fall outside procDomain
    c = sd.dambreak_case(nf=8)
    o = oracle.Oracle(c)
    assert o.num_deleted() > 0
    assert o.num_particles() == c.n - o.num_deleted()
    r = o.get("r")
    alive = ~np.isnan(r[:, 0])
    lo = np.array(c.params.proc_domain[0::2])
    assert np.all(r[alive] >= lo - 1e-12)


def test_periodic_wrap_reenters():
    c = sd.sediment_case(nf=12, n_dem=0)
    o = oracle.Oracle(c)
    r = c.r.copy()
    r[0, 0] = 1.0 + 0.3 / 12  # just past RMAX in periodic x
    o.set("r", r)
    o.reset()
    got = o.get("r")
    assert got[0, 0] == pytest.approx(r[0, 0] - 1.0, abs=1e-15)
    assert o.num_deleted() == 0


# ------------------------------------------------------- lattice sums
def periodic_lattice_case(n=12):
    """A perfect cubic lattice filling a fully periodic unit box."""
    c = sd.sediment_case(nf=n, n_dem=0, jitter=False)
    sph = c.iam == sd.SPH
    P = c.params.copy()
    P.periodic[2] = 1
    for i in range(3):
        P.proc_domain[2 * i] = 0.0
        P.proc_domain[2 * i + 1] = 1.0
    P.flags = sd.F_VISC_MORRIS | sd.F_VAR_H_CORRECTION
    return sd.Case("lattice", P, c.r[sph], c.v[sph], c.vhat[sph], c.mass[sph], c.h[sph],
                   c.dens[sph], c.iam[sph], sd.HOOKS_NONE)


def lattice_sum(fn, dx, hmax):
    m = int(math.ceil(2 * hmax / dx)) + 1
    g = np.arange(-m, m + 1)
    i, j, k = np.meshgrid(g, g, g, indexing="ij")
    d = np.sqrt((i * dx) ** 2 + (j * dx) ** 2 + (k * dx) ** 2).ravel()
    d = d[d > 0]
    return fn(d)


def test_density_and_omega_on_lattice():
    c = periodic_lattice_case(12)
    o = oracle.Oracle(c)
    dx = c.params.psep
    h = c.h[0]
    m = c.mass[0]
    o.pass_("density")  # one calcHFromDens + calcDensity iteration
    h1 = o.get("h")[0]
    rho = o.get("dens")
    wcon = c.params.wcon

    def W(d, hh):
        q = d / hh
        return np.where(q <= 2, wcon * (2 - q) ** 4 * (1 + 2 * q) / hh ** 3, 0.0)

    exp = m * W(np.array([0.0]), h1)[0] + m * np.sum(lattice_sum(lambda d: W(d, h1), dx, h1))
    assert np.allclose(rho, exp, rtol=1e-13)
    o2 = oracle.Oracle(c)
    o2.reset()
    o2.pass_("omega")
    om = o2.get("omega")

    def pdWh(d, hh):
        q = d / hh
        K = np.where(q <= 2, wcon * (2 - q) ** 4 * (1 + 2 * q), 0.0)
        pdK = np.where(q <= 2, -wcon * q * (-4 * (2 - q) ** 3 * (1 + 2 * q) + 2 * (2 - q) ** 4) / hh,
                       0.0)
        return pdK / hh ** 3 - 3 / hh ** 4 * K

    s = m * pdWh(np.array([0.0]), h)[0] + m * np.sum(lattice_sum(lambda d: pdWh(d, h), dx, h))
    exp_om = 1 + h * s / (c.dens[0] * 3)
    assert np.allclose(om, exp_om, rtol=1e-12)


def test_uniform_lattice_has_no_pressure_force():
    c = periodic_lattice_case(12)
    o = oracle.Oracle(c)
    o.reset()
    o.pass_("omega")
    o.pass_("eos")
    d = o.get("dens") * 1.01  # nonzero pressure
    o.set("dens", d)
    o.refresh_ghosts()
    o.pass_("eos")
    o.pass_("force")
    f = o.get("f")
    p = o.get("pdr2")
    scale = c.mass[0] * np.abs(p).max() * 2 * abs(o.F(1.0, c.h[0])) * c.h[0]
    assert np.abs(f).max() < 1e-12 * scale * 100


def test_pair_forces_conserve_momentum():
    """Without VAR_H the pressure and Morris forces are antisymmetric per pair:
    sum_a m_a f_a = 0 in a periodic box."""
    c = periodic_lattice_case(12)
    P = c.params.copy()
    P.flags = sd.F_VISC_MORRIS | sd.F_VISC_ARTIFICIAL
    c = sd.Case("lat", P, sd._jitter(c.r, P.psep, 5, 0.05) % 1.0, c.v, c.vhat, c.mass, c.h,
                c.dens, c.iam, sd.HOOKS_NONE)
    o = oracle.Oracle(c)
    rng = np.random.default_rng(0)
    o.set("vhat", rng.normal(size=(c.n, 3)) * 0.1)
    o.set("dens", c.dens * (1 + 0.01 * rng.random(c.n)))
    o.refresh_ghosts()
    o.pass_("eos")
    o.pass_("force")
    f = o.get("f")
    mom = (c.mass[:, None] * f).sum(0)
    assert np.abs(mom).max() < 1e-10 * (c.mass[:, None] * np.abs(f)).sum()
    assert np.abs(f).max() > 0


@pytest.mark.parametrize("visc", ["cleary", "monaghan"])
def test_viscosity_variants_conserve_dissipate_and_ignore_rotation(visc):
    """Cleary and Monaghan viscosity (sph.h:598-630) act along dr with a
    coefficient in dv.dr: per pair antisymmetric (sum_a m_a fv_a = 0 in a
    periodic box without VAR_H), dissipative (sum_a m_a vhat_a.fv_a < 0, since
    the gradient factor F < 0), and zero where dv.dr = 0 for every pair (a
    rigid rotation), which Morris viscosity is not."""
    flag = {"cleary": sd.F_VISC_CLEARY, "monaghan": sd.F_VISC_MONAGHAN}[visc]
    c = periodic_lattice_case(12)
    P = c.params.copy()
    P.flags = flag
    c = sd.Case("lat", P, sd._jitter(c.r, P.psep, 5, 0.05) % 1.0, c.v, c.vhat, c.mass, c.h,
                c.dens, c.iam, sd.HOOKS_NONE)
    o = oracle.Oracle(c)
    rng = np.random.default_rng(1)
    vh = rng.normal(size=(c.n, 3)) * 0.1
    o.set("vhat", vh)
    o.set("dens", c.dens * (1 + 0.01 * rng.random(c.n)))
    o.refresh_ghosts()
    o.pass_("eos")
    o.pass_("force")
    fv = o.get("fv")
    m = c.mass[:, None]
    assert np.abs(fv).max() > 0
    assert np.abs((m * fv).sum(0)).max() < 1e-10 * (m * np.abs(fv)).sum()
    assert (m[:, 0] * (vh * fv).sum(1)).sum() < 0
    # rigid rotation about the box centre: interior particles (all partners
    # real, no periodic images) feel no viscous force; Morris's does not vanish
    w = np.array([0.3, -0.2, 0.5])
    rot = np.cross(w, c.r - 0.5)
    inner = np.all(np.abs(c.r - 0.5) < 0.5 - 3 * P.h, axis=1)
    scale = np.abs(fv).max()
    for flags, vanishes in ((flag, True), (sd.F_VISC_MORRIS, False)):
        P2 = P.copy()
        P2.flags = flags
        o2 = oracle.Oracle(c, params=P2)
        o2.set("vhat", rot)
        o2.refresh_ghosts()
        o2.pass_("eos")
        o2.pass_("force")
        fr = np.abs(o2.get("fv")[inner]).max()
        if vanishes:
            assert fr < 1e-9 * scale
        else:
            assert fr > 1e-3 * scale


def _contact_sediment():
    c = sd.sediment_case(nf=12, n_dem=40, seed=77)
    dem = np.where(c.iam == sd.DEM)[0]
    R = c.params.dem_radius
    for k in range(0, len(dem) - 1, 2):
        c.r[dem[k + 1]] = c.r[dem[k]] + np.array([1.9 * R, 0.0, 0.0])
    return c


def test_const_h_keeps_h():
    """CONST_H: driftRest leaves h alone (sphIncompress.h:699)."""
    c = sd.dambreak_case(nf=8)
    c.params.flags |= sd.F_CONST_H
    o = oracle.Oracle(c)
    o.run(3)
    alive = o.get("iam") != -1
    assert np.array_equal(o.get("h")[alive], c.h[alive])
    assert not np.array_equal(o.get("dens")[alive], c.dens[alive])


def test_damptime_scales_the_first_kick():
    """While time < DAMPTIME every kicked velocity is multiplied by 0.98
    (sphIncompress.h:716-721): after the first step (dt = 0 in start(), so the
    only kick is middle()'s) the damped SPH velocities are 0.98 times the
    undamped ones, bit for bit."""
    c = sd.dambreak_case(nf=8)
    a = oracle.Oracle(c)
    P = c.params.copy()
    P.damptime = 1
    b = oracle.Oracle(c, params=P)
    a.run(1)
    b.run(1)
    sph = c.iam == sd.SPH
    va, vb = a.get("v")[sph], b.get("v")[sph]
    assert np.abs(va).max() > 0
    assert np.array_equal(vb, va * 0.98)


def test_fixed_dem_never_moves_and_feels_no_contacts():
    """FIXED_DEM: DEM velocities are restored after each kick and contacts
    are skipped (sphIncompress.h:666-671, 737-741; sphIncompress.cpp:229)."""
    c = _contact_sediment()
    c.params.flags |= sd.F_FIXED_DEM
    o = oracle.Oracle(c)
    dem = c.iam == sd.DEM
    o.run(1)  # the first reset wraps spheres placed past the periodic faces
    r1 = o.get("r")[dem]
    o.run(2)
    assert np.array_equal(o.get("v")[dem], np.zeros((dem.sum(), 3)))
    assert np.array_equal(o.get("r")[dem], r1)


def test_no_dem_contacts_skips_the_contact_force():
    """NO_DEM_CONTACTS: overlapping spheres get no contact force (fp stays
    zero, while the default run pushes them apart)."""
    c = _contact_sediment()
    dem = c.iam == sd.DEM
    P = c.params.copy()
    P.flags |= sd.F_NO_DEM_CONTACTS
    with_c, without = oracle.Oracle(c), oracle.Oracle(c, params=P)
    with_c.run(2)
    without.run(2)
    assert np.abs(with_c.get("fp")[dem]).max() > 0
    assert np.abs(without.get("fp")[dem]).max() == 0


def test_dddt_vanishes_under_rigid_translation(sed):
    o = oracle.Oracle(sed)
    o.set("vhat", np.tile([0.3, -0.2, 0.1], (sed.n, 1)))
    o.refresh_ghosts()
    o.pass_("omega")
    o.pass_("dddt")
    dddt = o.get("dddt")
    assert np.abs(dddt[sed.iam != sd.DEM]).max() < 1e-9


def test_eos_closed_form(dam):
    o = oracle.Oracle(dam)
    rho = dam.params.refd * (1 + np.linspace(-0.01, 0.01, dam.n))
    o.set("dens", rho)
    o.pass_("eos")
    p = o.get("press")
    exp = dam.params.prb * ((rho / dam.params.refd) ** 7 - 1)
    assert np.allclose(p, exp, rtol=1e-12, atol=1e-9)
    assert np.allclose(o.get("pdr2"), p / rho ** 2, rtol=1e-14)


# ------------------------------------------------------- integrator
def single_particle_case(dt):
    c = sd.dambreak_case(nf=12)
    P = c.params.copy()
    P.maxdt = dt
    r = np.array([[0.5 * P.rmax[0], 0.5 * P.rmax[1], 0.5 * P.rmax[2]]])
    return sd.Case("one", P, r, np.zeros((1, 3)), np.zeros((1, 3)), c.mass[:1], c.h[:1],
                   c.dens[:1], np.array([sd.SPH], np.int32), sd.HOOKS_DAMBREAK)


def test_free_fall_trajectory():
    """One sph particle under gravity g with dt fixed by maxdt = D: the first
    step runs start() with dt = 0 (globalVars.h:16) and the kick uses
    dt_old/2 + dt/2, so after step n: z - z0 = g D^2 (2n^2 - 2n + 1)/4,
    vz = g D (n - 1/2), time = (n - 1/2) D."""
    D = 1e-3
    c = single_particle_case(D)
    o = oracle.Oracle(c)
    z0 = c.r[0, 2]
    g = -9.81
    for n in range(1, 6):
        o.run(1)
        z = o.get("r")[0, 2]
        assert z - z0 == pytest.approx(g * D * D * (2 * n * n - 2 * n + 1) / 4, rel=1e-12)
        assert o.get("vhat")[0, 2] == pytest.approx(g * D * (n - 0.5), rel=1e-12)
        assert o.globals().time == pytest.approx((n - 0.5) * D, rel=1e-12)


def test_dt_rule_courant_and_accel(dam):
    o = oracle.Oracle(dam)
    o.run(1)
    P = dam.params
    dt = o.globals().dt
    h = dam.h
    courant = 0.6 * h.min() / (2 * P.spsound)
    # the pair Courant condition with hav >= min h bounds dt; accel 0.25 sqrt(h/|f|)
    assert dt <= 0.6 * h.max() / (2 * P.spsound) * (1 + 1e-12)
    assert dt >= courant * 0.999
    assert dt <= 0.25 * math.sqrt(h.max() / 9.81)


# ------------------------------------------------------- DEM and coupling
def test_di_felice_drag_closed_form(sed):
    o = oracle.Oracle(sed)
    P = sed.params
    R = P.dem_radius
    m = P.dem_vol * P.dem_dens
    for eps, u in ((0.9, 0.01), (0.6, 0.2), (0.45, 1e-4)):
        fvel = np.array([0.0, 0.0, u])
        vhat = np.array([0.0, 0.0, 0.0])
        got = o.particle_drag(eps, vhat, fvel, m)
        # Di Felice (1994): F = 1/2 rho Cd pi R^2 |U| U eps^(-chi), U = eps (u_f - u_p),
        # Re = 2 R |U| / nu, Cd = (0.63 + 4.8 Re^-1/2)^2,
        # chi = 3.7 - 0.65 exp(-(1.5 - log10 Re)^2 / 2); acceleration = F / m
        U = eps * u
        Re = 2 * R * U / P.viscosity
        Cd = (0.63 + 4.8 / math.sqrt(Re)) ** 2
        chi = 3.7 - 0.65 * math.exp(-((1.5 - math.log10(Re)) ** 2) / 2)
        F = 0.5 * P.dens * Cd * sd.PI_REF * R * R * U * U * eps ** (-chi)
        assert got[2] == pytest.approx(F / m, rel=1e-12)
        assert got[0] == 0.0 and got[1] == 0.0


def dem_pair_case(v_rel=0.05):
    c = sd.sediment_case(nf=12, n_dem=0)
    P = c.params.copy()
    # without fluid nothing bounds the first dt (DEM_TIMESTEP applies only once
    # time > TIME_DROP_PARTICLE, sphIncompress.cpp:133), so cap it by maxdt
    P.maxdt = P.dem_timestep
    R = P.dem_radius
    r = np.array([[0.5 - 1.05 * R, 0.5, 0.5], [0.5 + 1.05 * R, 0.5, 0.5]])
    v = np.array([[v_rel / 2, 0, 0], [-v_rel / 2, 0, 0]])
    m = P.dem_vol * P.dem_dens
    return sd.Case("pair", P, r, v.copy(), v.copy(), np.full(2, m), np.full(2, R),
                   np.full(2, P.dem_dens), np.full(2, sd.DEM, np.int32), sd.HOOKS_NONE)


def test_two_sphere_restitution():
    """Head-on collision of two DEM spheres with the linear spring-dashpot:
    restitution e = exp(-beta t_c), beta = gamma/(2 m_red), t_c = pi/sqrt(k/m_red - beta^2).
    The constants were built for e = 0.9."""
    c = dem_pair_case(0.05)
    o = oracle.Oracle(c)
    for _ in range(400):
        o.run(1)
        v = o.get("v")
        if v[0, 0] < 0 and v[1, 0] > 0:
            r = o.get("r")
            if r[1, 0] - r[0, 0] > 2 * c.params.dem_radius:
                break
    v = o.get("v")
    e = (v[1, 0] - v[0, 0]) / 0.05
    P = c.params
    m_red = 0.5 * P.dem_vol * P.dem_dens
    beta = P.dem_gamma / (2 * m_red)
    tc = math.pi / math.sqrt(P.dem_k / m_red - beta ** 2)
    assert e == pytest.approx(math.exp(-beta * tc), rel=0.02)
    assert e == pytest.approx(0.9, rel=0.02)
    # momentum is conserved exactly up to round-off (equal masses)
    assert abs(v[0, 0] + v[1, 0]) < 1e-15


def test_porosity_and_shepard_sum_consistency(sed):
    """sum_a V_a (1 - eps_a) = (4/3 pi R^3) sum_j S_j before clamping: the
    porosity gather and the Shepard scatter see the same pair set."""
    o = oracle.Oracle(sed)
    o.pass_("porosity")
    eps = o.get("porosity")
    S = o.get("shepSum")
    sph = sed.iam != sd.DEM
    dem = ~sph
    assert np.all(eps[sph] > 0.37)  # no clamping in this case
    V = sed.mass[sph] / sed.dens[sph]
    lhs = np.sum(V * (1 - eps[sph]))
    rhs = sed.params.porosity_incr * np.sum(S[dem])
    assert lhs == pytest.approx(rhs, rel=1e-12)
    # Shepard sums of DEM inside the fluid are close to a partition of unity
    assert np.all(np.abs(S[dem] - 1) < 0.05)


def test_dem_to_sph_drag_is_newtons_third_law(sed):
    """sum_a m_a fdrag_a = - sum_j m_j fdrag_j (calcDEMtoSPHDrag divides by the
    Shepard sum of exactly the same pairs)."""
    o = oracle.Oracle(sed)
    rng = np.random.default_rng(1)
    fd = np.zeros((sed.n, 3))
    dem = sed.iam == sd.DEM
    fd[dem] = rng.normal(size=(dem.sum(), 3))
    o.set("fdrag", fd)
    o.refresh_ghosts()
    o.pass_("porosity")
    fd2 = o.get("fdrag")
    sph = ~dem
    lhs = (sed.mass[sph, None] * fd2[sph]).sum(0)
    rhs = -(sed.mass[dem, None] * fd[dem]).sum(0)
    assert np.allclose(lhs, rhs, rtol=1e-11, atol=1e-14)


def test_sediment_run_is_finite_and_dt_is_bounded(sed):
    o = oracle.Oracle(sed)
    o.run(3)
    for name in ("r", "v", "f", "dens", "h", "porosity"):
        x = o.get(name)
        assert np.all(np.isfinite(x)), name
    P = sed.params
    assert o.globals().dt <= P.dem_timestep * (1 + 1e-15)


def test_parallel_oracle_equals_the_serial_sweep():
    """The oracle's OpenMP gathers leave every result bit for bit as the
    one-thread sweep leaves it (each particle's sums keep their order; the dt
    minimum is exact in any order)."""
    c = sd.sediment_case(nf=10, n_dem=40, seed=3)
    out = {}
    for nt in (1, 4):
        oracle.set_threads(nt)
        o = oracle.Oracle(c)
        o.run(2)
        out[nt] = {k: o.get(k) for k in ("r", "v", "vhat", "dens", "h", "f", "porosity",
                                        "fdrag", "shepSum", "omega", "dddt")}
        out[nt]["dt"] = o.globals().dt
        o.close()
    oracle.set_threads(1)
    for k in out[1]:
        assert np.array_equal(out[1][k], out[4][k], equal_nan=True), k


def test_dem_block_gives_every_sphere_the_mean_force():
    """DEM_BLOCK (sphIncompress.cpp:231-238): after end() every DEM particle
    carries the mean of the forces the DEM particles had before it, and the
    kick used that mean (v = v0 + dt f)."""
    c = _contact_sediment()
    dem = c.iam == sd.DEM
    P = c.params.copy()
    P.flags |= sd.F_DEM_BLOCK
    plain, block = oracle.Oracle(c), oracle.Oracle(c)
    for o in (plain, block):
        o.run(1)
        if o is block:
            o.set_params(P)  # the same state, then one step with the block
        g = o.globals()
        g.sphStep += 1
        o.set_globals(g)
        o.start()
        o.middle()
        o.end()
    fb, fp = block.get("f")[dem], plain.get("f")[dem]
    assert np.all(fb == fb[0])
    assert np.allclose(fb[0], fp.mean(axis=0), rtol=1e-12, atol=1e-15)
    dt = block.globals().dt
    assert np.allclose(block.get("v")[dem], block.get("v0")[dem] + dt * fb, rtol=1e-12,
                       atol=1e-15)


def test_wall_contact_scan_of_the_dem_lists_changes_nothing():
    """Without LIQ_DEM_CUSTOM_WALL_CONTACTS the reference adds
    calcDEMWallContacts (sph.h:1634-1674), which looks for wall particles in
    the DEM's own neighbour list; DEM lists hold DEM only, so the run is the
    same bit for bit."""
    c = _contact_sediment()
    P = c.params.copy()
    P.flags &= ~sd.F_CUSTOM_WALL
    a, b = oracle.Oracle(c), oracle.Oracle(c, params=P)
    a.run(3)
    b.run(3)
    for k in ("r", "v", "f", "fp", "fb", "dens"):
        assert np.array_equal(a.get(k), b.get(k), equal_nan=True), k
