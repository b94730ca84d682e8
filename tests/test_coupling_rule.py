"""Known answers of calcNeighboursCoupling's SPH-side cell rule
(dataLL.cpp:1154-1190), worked by hand from the reference's text, for the
oracle and (on the GPU) the library.

The rule, for an SPH particle in DEM cell c_a and the DEM cells c_a + d with
|d_i| <= nc = ceil(0.5 Rc / DSR): with plus = sum_i (d_i + 1)^2, minus =
sum_i (d_i - 1)^2 and t^2 = (0.5 Rc / DSR)^2, a cell with plus < t^2 is taken
whole (addAll, whatever the distance); otherwise a cell with minus < t^2 is
searched for DEM particles within Rc (addIfInRadius); other cells are never
looked at.  The +1 on every axis of `plus` and the -1 on every axis of
`minus` make the rule lopsided: towards negative offsets it takes whole cells
beyond Rc, and at mixed-sign offsets it skips cells that hold DEM particles
within Rc.  Each case below puts one DEM particle at a chosen spot and states
the answer this text gives; a closest-corner rule, swapped plus and minus,
<= for <, or a sign flip of d each fail at least one of them.

Geometry: the sediment constants, Rc = 0.25 and DSR = 1/48, so a DEM cell
edge is e = 2 DSR = Rc / 6, t^2 = 36 and nc = 6.  DEM cell boundaries lie at
cmin + m e (the cell formula (x - cmin)/e + 2 Rc/DSR - 1 shifts by a whole
23).  The SPH particle sits in the cell [0.5, 0.5 + e)^3 (x, y from cmin = 0,
z from cmin = -0.25, 18 cells up), far from every periodic face.
"""
import dataclasses
import math

import numpy as np
import pytest

import oracle
import sphdem as sd

BASE = np.array([0.5, 0.5, 0.5])


def geometry():
    P = sd.sediment_case(nf=12, n_dem=60).params
    e = 2.0 * P.dem_search_radius
    assert math.isclose(P.coupling_radius, 6.0 * e, rel_tol=1e-12)
    assert math.ceil(0.5 * P.coupling_radius / P.dem_search_radius) == 6
    # BASE is a DEM cell corner on every axis
    for i in range(3):
        m = (BASE[i] - P.proc_domain[2 * i]) / e
        assert abs(m - round(m)) < 1e-9
    return P, e


# (name, SPH position in its cell (units of e), DEM cell offset d, DEM
#  position in that cell (units of e), coupled?, distance / Rc)
CASES = [
    # plus = 25+1+1 = 27 < 36: the whole cell, although the particle is
    # 6.48 e > Rc away
    ("whole cell beyond Rc, -x", (0.5, 0.5, 0.5), (-6, 0, 0), (0.02, 0.5, 0.5), True, 1.08),
    # the mirror image: plus = 51, minus = 27: searched, 6.48 e > Rc
    ("searched cell beyond Rc, +x", (0.5, 0.5, 0.5), (6, 0, 0), (0.98, 0.5, 0.5), False, 1.08),
    # the same cell at 5.52 e < Rc: found by the search
    ("searched cell within Rc, +x", (0.5, 0.5, 0.5), (6, 0, 0), (0.02, 0.5, 0.5), True, 0.92),
    # plus = 36+9+1 = 46, minus = 16+25+1 = 42: never looked at, although
    # the particle is 5.73 e < Rc away
    ("skipped cell within Rc, (+5,-4)", (0.5, 0.5, 0.5), (5, -4, 0), (0.02, 0.98, 0.5), False,
     0.955),
    ("skipped cell within Rc, (-4,+5)", (0.5, 0.5, 0.5), (-4, 5, 0), (0.98, 0.02, 0.5), False,
     0.955),
    # plus = 25+9+1 = 35 < 36: whole, the far corner 6.34 e > Rc away
    ("whole cell beyond Rc, (+4,-4)", (0.5, 0.5, 0.5), (4, -4, 0), (0.98, 0.02, 0.5), True, 1.056),
    # plus = 4+4+4 = 12: whole, the far corner 6.03 e > Rc away
    ("whole cell beyond Rc, (-3,-3,-3)", (0.5, 0.5, 0.5), (-3, -3, -3), (0.02, 0.02, 0.02), True,
     1.005),
    # plus = 48, minus = 12: searched; far corner 6.03 e > Rc, near corner
    # 4.36 e
    ("searched cell beyond Rc, (+3,+3,+3)", (0.5, 0.5, 0.5), (3, 3, 3), (0.98, 0.98, 0.98), False,
     1.005),
    ("searched cell within Rc, (+3,+3,+3)", (0.5, 0.5, 0.5), (3, 3, 3), (0.02, 0.02, 0.02), True,
     0.727),
    # plus = 36+0+0 = 36, not < 36, so searched (minus = 16+4+4 = 24):
    # 6.13 e > Rc, not coupled (<= in place of < would take it whole)
    ("plus equal to t^2, (+5,-1,-1)", (0.02, 0.02, 0.02), (5, -1, -1), (0.98, 0.02, 0.02), False,
     1.021),
]


def one_case(P, e, sph, d, dem):
    r = np.array([BASE + e * np.array(sph),
                  BASE + e * (np.array(d, dtype=float) + np.array(dem))])
    n = 2
    iam = np.array([sd.SPH, sd.DEM], dtype=np.int32)
    mass = np.array([P.dens * (1.0 / 12) ** 3, 2500.0 * P.dem_vol])
    case = sd.Case(name="coupling-rule", params=P.copy(), r=r, v=np.zeros((n, 3)),
                   vhat=np.zeros((n, 3)), mass=mass, h=np.full(n, P.h),
                   dens=np.full(n, P.dens), iam=iam, hookset=sd.HOOKS_NONE)
    return case, float(np.linalg.norm(r[0] - r[1]))


@pytest.mark.parametrize("name,sph,d,dem,coupled,dist", CASES, ids=[c[0] for c in CASES])
def test_oracle_coupling_rule_known_answers(name, sph, d, dem, coupled, dist):
    P, e = geometry()
    case, l = one_case(P, e, sph, d, dem)
    assert l / P.coupling_radius == pytest.approx(dist, abs=2e-3)
    got = oracle.Oracle(case).pairs(sd.PAIRS_SPH2DEM)
    assert (len(got) == 1) == coupled, got


@pytest.mark.gpu
def test_library_coupling_rule_known_answers():
    P, e = geometry()
    for name, sph, d, dem, coupled, dist in CASES:
        case, _ = one_case(P, e, sph, d, dem)
        sim = sd.Sim(case)
        sim.reset()
        got = sim.pairs(sd.PAIRS_SPH2DEM)
        sim.close()
        assert (len(got) == 1) == coupled, name
