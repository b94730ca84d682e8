"""calcOutputVars diagnostics (sphdem_output_vars) against the same sums
restated over the downloaded particle fields (sphIncompress.h:233-319)."""
import numpy as np
import pytest

import sphdem as sd

pytestmark = pytest.mark.gpu


def test_output_vars_match_field_sums():
    c = sd.sediment_case(nf=12, n_dem=60)
    sim = sd.Sim(c)
    sim.run(1)
    o = sim.output_vars()
    iam = sim.get("iam")
    sph = iam == sd.SPH
    grp = sph | (iam == sd.SPH_BOUNDARY)
    n = int(np.sum(iam != sd.DELETED))
    m, v, vh, r = sim.get("mass"), sim.get("v"), sim.get("vhat"), sim.get("r")
    dens, f, fv = sim.get("dens"), sim.get("f"), sim.get("fv")
    u, dudt = sim.get("u"), sim.get("dudt")
    g = sim.globals()
    P = c.params

    def dot(a, b):  # blitz association
        return a[:, 0] * b[:, 0] + (a[:, 1] * b[:, 1] + a[:, 2] * b[:, 2])

    dedt = -dot(vh, fv)
    df = dens / P.refd
    want = {
        "eKE": np.sum(0.5 * m[sph] * dot(vh, vh)[sph]),
        "angMom": np.sum(((r[:, 0] * v[:, 1] - r[:, 1] * v[:, 0]) * m)[sph]),
        "edViscFdt": np.sum((m * dedt)[sph]),
        # one step: each sph particle's running energy is dt * dE/dt
        "eViscF": np.sum((m * (g.dt * dedt))[sph]),
        "eElast": np.sum((m * u)[grp]),
        "edElastdt": np.sum((m * dudt)[grp]),
        "eElastExact": np.sum((m * (P.prb * ((df ** 6 - 1) / 6 + 1 / df - 1)) / P.refd)[grp]),
        "aveDens": np.sum(dens[grp]) / n,
        "maxV": np.max(np.sqrt(dot(v, v))[grp]),
        "maxF": np.max(np.sqrt(dot(f, f))[grp]),
    }
    want["varDens"] = np.sum((dens[grp] - want["aveDens"]) ** 2) / n
    for k, w in want.items():
        assert o[k] == pytest.approx(w, rel=1e-12, abs=1e-300), k
    assert o["linMom"] == pytest.approx(list(np.sum((v * m[:, None])[sph], axis=0)), rel=1e-12)
    assert o["n"] == n and o["nSph"] == int(sph.sum())
    assert o["eTotal"] == pytest.approx(o["eKE"] + o["eElast"] + o["eViscF"], rel=1e-12)


def test_density_from_mass_vanishes_after_a_density_pass():
    """Right after calcDensity over the cached lists, the fresh summation of
    calcAveDensFromMass reproduces every density exactly."""
    sim = sd.Sim(sd.sediment_case(nf=12, n_dem=60))
    sim.run(1)
    sim.reset()
    sim.pass_("density")
    o = sim.output_vars()
    assert o["aveDensFromMass"] == 0.0
    sim.run(1)
    assert sim.output_vars()["aveDensFromMass"] > 0.0


@pytest.mark.parametrize("case", ["sediment", "closedbox"])
def test_output_vars_match_the_oracle(case):
    """sphdem_output_vars after several steps against the oracle's
    calcOutputVars with its calcEnergies (sphIncompress.h:233-319, 769-781):
    the running viscous energy, the elastic energies, momenta, maxima and the
    density statistics, within the whole-step bound of the parity tests; the
    counts exactly."""
    import oracle

    c = (sd.sediment_case(nf=12, n_dem=60) if case == "sediment"
         else sd.dambreak_case(nf=12, with_dem=28, liq_dem_flags=True))
    sim, orc = sd.Sim(c), oracle.Oracle(c)
    sim.run(3)
    orc.run(3)
    g, o = sim.output_vars(), orc.output_vars()
    assert (g["n"], g["nSph"]) == (o["n"], o["nSph"])
    for k in ("angMom", "eKE", "eViscF", "edViscFdt", "eElast", "edElastdt", "eElastExact",
              "eTotal", "maxV", "maxF", "aveDens", "varDens", "aveDensFromMass"):
        assert g[k] == pytest.approx(o[k], rel=1e-9, abs=1e-12 * max(abs(o["eKE"]), 1.0)), k
    for i in range(3):
        assert g["linMom"][i] == pytest.approx(o["linMom"][i], rel=1e-9,
                                               abs=1e-12 * abs(o["linMom"][2])), i
    assert o["eViscF"] != 0.0 and o["eElast"] != 0.0
    for k in ("eViscB", "eBForce", "eFF", "edFFdt", "maxFF", "rmsFF"):
        assert g[k] == 0.0 and o[k] == 0.0, k
    sim.close()
    orc.close()
