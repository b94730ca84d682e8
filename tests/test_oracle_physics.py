"""Pins the oracle's pair operators by what the continuum mathematics fixes on
a perfect lattice, so that a wrong constant, sign or factor fails: the SPH
pressure force of a linear pressure field is -grad(p)/rho, the Morris
viscous force of a quadratic shear is nu lap(v), the continuity sum of a
uniform compression is 3 alpha rho, the LIQ_DEM equation of state vanishes at
rho = REFD * porosity with slope PRB gamma / (REFD porosity), and calcFPI in a
uniform fluid returns the Shepard-normalised fields in closed form (including
the reference's second fv term, sph.h:1744,1747).  Only interior particles
(all neighbours within the support present) are checked.  CPU only."""
import math

import numpy as np
import pytest

import oracle
import sphdem as sd

N = 14  # lattice particles per axis; dx = 1/N, h = 1.5 dx


def open_lattice(flags, n=N, dem_at=None):
    """A perfect cubic lattice of sph particles filling the unit cube, no
    walls and no periodic axes (the domain extends 0.25 beyond the cube), the
    constants of the sedimentation case; optionally one DEM sphere at
    `dem_at`."""
    base = sd.sediment_case(nf=n, n_dem=0, jitter=False)
    sph = base.iam == sd.SPH
    P = base.params.copy()
    for i in range(3):
        P.periodic[i] = 0
        P.rmin[i], P.rmax[i] = 0.0, 1.0
        P.proc_domain[2 * i], P.proc_domain[2 * i + 1] = -0.25, 1.25
    P.flags = flags
    r, mass, h, dens = base.r[sph], base.mass[sph], base.h[sph], base.dens[sph]
    iam = base.iam[sph]
    if dem_at is not None:
        r = np.vstack([r, np.asarray(dem_at, float)[None, :]])
        mass = np.append(mass, P.dem_vol * P.dem_dens)
        h = np.append(h, P.dem_search_radius)
        dens = np.append(dens, P.dem_dens)
        iam = np.append(iam, np.int32(sd.DEM)).astype(np.int32)
    z = np.zeros_like(r)
    return sd.Case("open_lattice", P, r, z.copy(), z.copy(), mass, h, dens, iam, sd.HOOKS_NONE)


def interior(c):
    """Particles farther than the support 2h (+ dx/4) from every face."""
    h = c.params.h
    m = np.all((c.r > 2 * h + 0.25 * c.params.psep) & (c.r < 1 - 2 * h - 0.25 * c.params.psep),
               axis=1)
    return m & (c.iam == sd.SPH)


VARIANTS = {"const_h": sd.F_VISC_MORRIS, "var_h": sd.F_VISC_MORRIS | sd.F_VAR_H_CORRECTION}


def omega_of(o, c):
    if c.params.flags & sd.F_VAR_H_CORRECTION:
        o.pass_("omega")
        return o.get("omega")
    return np.ones(c.n)


@pytest.mark.parametrize("variant", sorted(VARIANTS))
def test_pressure_force_of_a_linear_pressure_field(variant):
    """p = p0 + G x (densities set so that the Tait EOS gives it): the pressure
    force per unit mass is -G/rho along x and zero across (calcPressForce,
    sph.h:692-734, -sum m_b (p_a/rho_a^2 + p_b/rho_b^2) gradW).  Under
    VAR_H_CORRECTION gradW carries 1/omega_a, so omega_a f = -G/rho."""
    c = open_lattice(VARIANTS[variant])
    P = c.params
    x = c.r[:, 0]
    p0, G = 0.02 * P.prb, 0.01 * P.prb  # Pa and Pa per metre
    p = p0 + G * x
    rho = P.refd * (p / P.prb + 1.0) ** (1.0 / 7.0)
    o = oracle.Oracle(c)
    o.set("dens", rho)
    om = omega_of(o, c)
    o.pass_("eos")
    assert np.allclose(o.get("press"), p, rtol=1e-12)
    o.pass_("force")
    fp = o.get("fp")
    inn = interior(c)
    assert inn.sum() >= 200
    want = -G / rho[inn]
    got = fp[inn, 0] * om[inn]
    # the lattice sum of r F(r) reproduces the continuum's -3/rho to 1.1 %
    assert np.abs(got / want - 1).max() < 0.02, np.abs(got / want - 1).max()
    assert np.abs(fp[inn, 1:]).max() < 1e-6 * np.abs(want).max()
    # the viscous force is zero at rest
    assert np.abs(o.get("fv")).max() == 0.0


@pytest.mark.parametrize("liq", [False, True])
@pytest.mark.parametrize("variant", sorted(VARIANTS))
def test_morris_viscosity_of_a_quadratic_shear(variant, liq):
    """vhat_x = S z^2: the Morris viscous force per unit mass is nu lap(v) =
    2 nu S along x (calcViscForce, sph.h:579-597: m_b visc (dv . (dx . gradW))
    / (r^2 + 0.01 h^2) with visc = nu (rho_a + rho_b)/(rho_a rho_b), or
    2 nu DENS/(rho_a rho_b) under LIQ_DEM), times 1/omega_a under VAR_H."""
    flags = VARIANTS[variant] | (sd.F_LIQ_DEM if liq else 0)
    c = open_lattice(flags)
    P = c.params
    S = 3.0  # 1/(m s)
    vh = np.zeros((c.n, 3))
    vh[:, 0] = S * c.r[:, 2] ** 2
    o = oracle.Oracle(c)
    o.set("vhat", vh)
    om = omega_of(o, c)
    o.pass_("eos")
    o.pass_("force")
    fv = o.get("fv")
    inn = interior(c)
    want = 2.0 * P.viscosity * S
    got = fv[inn, 0] * om[inn]
    # the 0.01 h^2 regulariser and the lattice sum: 0.9 %
    assert np.abs(got / want - 1).max() < 0.02, np.abs(got / want - 1).max()
    assert np.abs(fv[inn, 1:]).max() < 1e-9 * want
    # a linear shear has no Laplacian: zero to round-off on the lattice
    vh[:, 0] = S * c.r[:, 2]
    o2 = oracle.Oracle(c)
    o2.set("vhat", vh)
    omega_of(o2, c)
    o2.pass_("eos")
    o2.pass_("force")
    assert np.abs(o2.get("fv")[inn]).max() < 1e-9 * want


@pytest.mark.parametrize("variant", sorted(VARIANTS))
def test_continuity_of_a_uniform_compression(variant):
    """vhat = -alpha (r - r_c): drho/dt = -rho div(v) = 3 alpha rho
    (calcDddtDudt, sphIncompress.h:783-851: sum m_b (vhat_a - vhat_b).gradW_a,
    with gradW_a / omega_a under VAR_H)."""
    c = open_lattice(VARIANTS[variant])
    alpha = 0.7
    vh = -alpha * (c.r - 0.5)
    o = oracle.Oracle(c)
    o.set("vhat", vh)
    om = omega_of(o, c)
    o.pass_("dddt")
    dddt = o.get("dddt")
    inn = interior(c)
    want = 3.0 * alpha * c.dens[inn]
    got = dddt[inn] * om[inn]
    # 0.06 % on this lattice: tight enough that leaving out 1/omega_a (which
    # moves it by 0.25 %, to -0.19 %) fails
    assert np.abs(got / want - 1).max() < 0.001, np.abs(got / want - 1).max()
    # a rigid translation and a rigid rotation leave the density alone
    w = np.array([0.3, -0.2, 0.5])
    for v in (np.tile([0.4, -0.1, 0.2], (c.n, 1)), np.cross(w, c.r - 0.5)):
        o2 = oracle.Oracle(c)
        o2.set("vhat", v)
        omega_of(o2, c)
        o2.pass_("dddt")
        assert np.abs(o2.get("dddt")[inn]).max() < 1e-9 * want.max()


def test_liq_dem_eos_zero_and_slope():
    """Under LIQ_DEM the Tait EOS takes the fluid fraction: p =
    PRB ((rho / (REFD eps))^GAMMA - 1) (sphIncompress.h:871), so p = 0 exactly
    at rho = REFD eps and dp/drho there is PRB GAMMA / (REFD eps); GAMMA, not
    the fixed 7 of the plain form, is the exponent."""
    c = open_lattice(sd.F_VISC_MORRIS | sd.F_LIQ_DEM)
    P = c.params.copy()
    P.gamma = 5.0  # distinguishable from the plain form's fixed 7
    o = oracle.Oracle(c, params=P)
    n = c.n
    eps = 0.45 + 0.5 * np.linspace(0.0, 1.0, n)
    o.set("porosity", eps)
    o.set("dens", P.refd * eps)
    o.pass_("eos")
    assert np.abs(o.get("press")).max() <= 1e-9 * P.prb
    d = 1e-6
    slope = []
    for s in (+1, -1):
        o.set("dens", P.refd * eps * (1 + s * d))
        o.pass_("eos")
        slope.append(o.get("press"))
    dpdrho = (slope[0] - slope[1]) / (2 * d * P.refd * eps)
    assert np.allclose(dpdrho, P.prb * P.gamma / (P.refd * eps), rtol=1e-6)
    # pdr2 = p / rho^2
    rho = o.get("dens")
    assert np.allclose(o.get("pdr2"), o.get("press") / rho ** 2, rtol=1e-14)


def test_fluid_particle_interaction_in_a_uniform_fluid():
    """A DEM sphere inside a uniform fluid (every fluid particle with the same
    density rho, vhat, porosity, fp, fb, fv, f, fdrag) moving with the fluid:
    calcFPI (sph.h:1708-1792) Shepard-normalises by S = sum (m/rho) W, so
    fvel = vhat, porosity = eps, fb (the stored fp + fb) = rho (fp + fb) and
    fv = (rho + 1) fv (the sum m W fv plus the second sum (m/rho) W fv of
    sph.h:1744,1747); with no relative velocity the Di Felice term is zero and
    fdrag = (V/m)(fv_out + fb_out), which is also added to f."""
    c = open_lattice(sd.F_VISC_MORRIS | sd.F_VAR_H_CORRECTION | sd.F_LIQ_DEM,
                     dem_at=(0.5 + 0.013, 0.5 - 0.021, 0.5 + 0.007))
    P = c.params
    o = oracle.Oracle(c)
    n = c.n
    dem = c.iam == sd.DEM
    U = np.array([0.02, -0.01, 0.03])
    eps = 0.8
    fp, fb, fv = np.array([1.0, 2.0, -3.0]), np.array([-0.5, 0.25, 4.0]), np.array([0.3, -0.7, 0.1])
    fl = np.array([5.0, 6.0, 7.0])
    o.set("vhat", np.tile(U, (n, 1)))
    o.set("porosity", np.where(dem, 1.0, eps))
    for nm, v in (("fp", fp), ("fb", fb), ("fv", fv), ("f", fl), ("fdrag", np.zeros(3))):
        a = np.tile(v, (n, 1))
        a[dem] = 0.0
        o.set(nm, a)
    o.pass_("fpi")
    rho = c.dens[~dem][0]
    S = o.get("shepSum")[dem][0]
    assert abs(S - 1.0) < 0.02  # a partition of unity in the interior
    got = {k: o.get(k)[dem][0] for k in ("fvel", "porosity", "fb", "fv", "fdrag", "f")}
    assert np.allclose(got["fvel"], U, rtol=1e-12)
    assert got["porosity"] == pytest.approx(eps, rel=1e-12)
    assert np.allclose(got["fb"], rho * (fp + fb), rtol=1e-12)
    assert np.allclose(got["fv"], (rho + 1.0) * fv, rtol=1e-12)
    m_p = P.dem_vol * P.dem_dens
    want = (P.dem_vol / m_p) * ((rho + 1.0) * fv + rho * (fp + fb))
    assert np.allclose(got["fdrag"], want, rtol=1e-12)
    assert np.allclose(got["f"], want, rtol=1e-12)  # f started at zero


def test_fluid_particle_interaction_adds_the_di_felice_drag():
    """The same sphere at rest in a moving uniform fluid: fdrag gains the Di
    Felice drag of the relative velocity eps (U - 0) (Di Felice 1994, with
    the reference's PI = 3.14159265 in the area), evaluated here from the
    published correlation."""
    c = open_lattice(sd.F_VISC_MORRIS | sd.F_VAR_H_CORRECTION | sd.F_LIQ_DEM,
                     dem_at=(0.5, 0.5, 0.5))
    P = c.params
    o = oracle.Oracle(c)
    n = c.n
    dem = c.iam == sd.DEM
    U = np.array([0.0, 0.0, 0.05])
    eps = 0.7
    vh = np.tile(U, (n, 1))
    vh[dem] = 0.0
    o.set("vhat", vh)
    o.set("porosity", np.where(dem, 1.0, eps))
    o.pass_("fpi")
    R, nu = P.dem_radius, P.viscosity
    m_p = P.dem_vol * P.dem_dens
    Urel = eps * U[2]
    Re = 2 * R * Urel / nu
    chi = 3.7 - 0.65 * math.exp(-((1.5 - math.log10(Re)) ** 2) / 2)
    Cd = (0.63 + 4.8 / math.sqrt(Re)) ** 2
    F = 0.5 * P.dens * Cd * sd.PI_REF * R * R * Urel * Urel * eps ** (-chi)
    got = o.get("fdrag")[dem][0]
    assert got[2] == pytest.approx(F / m_p, rel=1e-10)
    assert abs(got[0]) < 1e-12 * abs(got[2]) and abs(got[1]) < 1e-12 * abs(got[2])


def test_output_vars_of_a_rigidly_moving_compressed_lattice():
    """calcOutputVars (sphIncompress.h:233-319) on a lattice moving rigidly
    with v = U (vhat = 1.5 U), every sph density REFD (1 + d), and one DEM
    sphere moving faster: the momenta are those of the rigid motion of the
    fluid mass alone, the kinetic energy is the fluid's in vhat (calcOutput
    sums 0.5 m vhat^2), maxV is |U| (the DEM sphere is not an sph/sphBoundary
    particle), the density mean and variance divide by every particle of the
    container, and the exact elastic energy is m times the integral of p /
    rho^2 from REFD to rho (by quadrature here).  After a density pass the
    density-from-mass statistic is zero."""
    from scipy.integrate import quad

    c = open_lattice(sd.F_VISC_MORRIS, dem_at=(0.5 + 0.013, 0.5 - 0.021, 0.5 + 0.007))
    P = c.params
    n = c.n
    sph = c.iam == sd.SPH
    U = np.array([0.3, -0.2, 0.1])
    v = np.tile(U, (n, 1))
    v[~sph] = 10.0 * U
    d = 0.01
    dens = np.where(sph, P.refd * (1.0 + d), c.dens)
    o = oracle.Oracle(c)
    o.set("v", v)
    o.set("vhat", 1.5 * v)
    o.set("dens", dens)
    out = o.output_vars()
    m, r = c.mass[sph], c.r[sph]
    M = m.sum()
    assert np.allclose(out["linMom"], M * U, rtol=1e-12)
    assert out["angMom"] == pytest.approx(float((m * (r[:, 0] * U[1] - r[:, 1] * U[0])).sum()),
                                          rel=1e-11)
    assert out["eKE"] == pytest.approx(0.5 * M * 2.25 * float(U @ U), rel=1e-12)
    assert out["maxV"] == pytest.approx(float(np.linalg.norm(U)), rel=1e-14)
    assert out["n"] == n and out["nSph"] == n - 1
    rho = P.refd * (1.0 + d)
    ave = (n - 1) * rho / n
    assert out["aveDens"] == pytest.approx(ave, rel=1e-12)
    assert out["varDens"] == pytest.approx((n - 1) * (rho - ave) ** 2 / n, rel=1e-9)
    e_unit, _ = quad(lambda x: P.prb * ((x / P.refd) ** 7 - 1.0) / x ** 2, P.refd, rho,
                     epsabs=0.0, epsrel=1e-13)
    assert out["eElastExact"] == pytest.approx(M * e_unit, rel=1e-9)
    assert out["eTotal"] == pytest.approx(out["eKE"] + out["eElast"] + out["eViscF"]
                                          + out["eViscB"] + out["eBForce"] + out["eFF"], rel=1e-14)
    o.pass_("density")
    assert o.output_vars()["aveDensFromMass"] == 0.0


def test_var_h_pair_terms_are_divided_by_the_query_particles_omega():
    """Under VAR_H_CORRECTION gradW divides both kernel terms of a pair by
    pa.omega, the query particle's own omega (sph.h:331-332), not each by its
    side's.  omega_a times the pressure and viscous force of a is then a sum of
    terms antisymmetric in (a, b), so sum_a m_a omega_a (fp_a + fv_a) vanishes
    to rounding, while sum_a m_a (fp_a + fv_a) does not where omega varies;
    dividing each side by its own omega would conserve sum_a m_a f_a instead.
    An open, jittered box with a density wave (h and omega follow it) and a
    velocity field that shears and rotates separates the two."""
    base = sd.sediment_case(nf=12, n_dem=0, jitter=True, seed=17)
    sph = base.iam == sd.SPH
    P = base.params.copy()
    for i in range(3):
        P.periodic[i] = 0
        P.rmin[i], P.rmax[i] = 0.0, 1.0
        P.proc_domain[2 * i], P.proc_domain[2 * i + 1] = -0.25, 1.25
    P.flags = sd.F_VISC_MORRIS | sd.F_VAR_H_CORRECTION
    r, mass, iam = base.r[sph], base.mass[sph], base.iam[sph]
    x, y, z = r.T
    dens = P.refd * (1.0 + 0.03 * np.sin(2 * np.pi * x) * np.cos(2 * np.pi * z) + 0.02 * y)
    h = P.hfac * (mass / dens) ** (1.0 / 3.0)
    vhat = np.stack([0.1 * z ** 2 - 0.05 * y, 0.05 * x + 0.02 * z,
                     0.03 * np.sin(2 * np.pi * y)], axis=1)
    c = sd.Case("varh_box", P, r, vhat.copy(), vhat.copy(), mass, h, dens, iam, sd.HOOKS_NONE)
    o = oracle.Oracle(c)
    o.pass_("omega")
    om = o.get("omega")
    assert om.std() > 1e-3 * om.mean()
    o.pass_("eos")
    o.pass_("force")
    # the pressure and the viscous parts each on their own scale
    for name in ("fp", "fv"):
        f = o.get(name)
        w = (mass * om)[:, None] * f
        u = mass[:, None] * f
        s_omega = np.abs(w.sum(axis=0)).max() / np.abs(w).sum(axis=0).max()
        s_plain = np.abs(u.sum(axis=0)).max() / np.abs(u).sum(axis=0).max()
        assert s_omega < 1e-12, (name, s_omega)
        assert s_plain > 1e-6, (name, s_plain)
    o.close()
