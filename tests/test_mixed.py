"""The mixed-precision form (SPHDEM_MIXED): FP32 pair arithmetic over the
FP64 positions, lists, sums, integration and dt reduction.  Checked against
the FP64 form of the library (itself at parity with the oracle) on the same
inputs: pair sets identical; after one step, density within 1e-5 relative
(the target error) and forces within 1e-4 normwise; after three steps
positions within 1e-9 of the box (normwise)."""
import numpy as np
import pytest

import sphdem as sd

pytestmark = pytest.mark.gpu

def _monaghan():
    c = sd.dambreak_case(nf=12)
    c.params.flags = (c.params.flags & ~sd.F_VISC_MORRIS) | sd.F_VISC_MONAGHAN
    return c


CASES = {"dambreak": lambda: sd.dambreak_case(nf=12),
         "sediment": lambda: sd.sediment_case(nf=12, n_dem=60),
         "dambreak_monaghan": _monaghan,
         # the large mixed-precision workload's geometry (closed box, SPH-DEM) at 12^3
         "closedbox": lambda: sd.dambreak_case(nf=12, with_dem=28, liq_dem_flags=True)}


def pair_of_sims(mk):
    return sd.Sim(mk(), precision=sd.FP64), sd.Sim(mk(), precision=sd.MIXED)


def rel_max(a, b):
    return float(np.max(np.abs(a - b) / np.abs(b)))


def normwise(a, b):
    return float(np.abs(a - b).max() / np.abs(b).max())


@pytest.mark.parametrize("name", sorted(CASES))
def test_mixed_pair_sets_are_the_fp64_sets(name):
    s64, s32 = pair_of_sims(CASES[name])
    s64.run(1)
    s32.run(1)
    a = s64.pairs(sd.PAIRS_SPH)
    b = s32.pairs(sd.PAIRS_SPH)
    assert np.array_equal(a, b)


@pytest.mark.parametrize("name", sorted(CASES))
def test_mixed_one_step_error(name):
    s64, s32 = pair_of_sims(CASES[name])
    s64.run(1)
    s32.run(1)
    sph = s64.get("iam") == sd.SPH
    assert rel_max(s32.get("dens")[sph], s64.get("dens")[sph]) <= 1e-5
    assert normwise(s32.get("f")[sph], s64.get("f")[sph]) <= 1e-4
    if (s64.get("iam") == sd.DEM).any():
        dem = s64.get("iam") == sd.DEM
        assert normwise(s32.get("f")[dem], s64.get("f")[dem]) <= 1e-4


@pytest.mark.parametrize("name", sorted(CASES))
def test_mixed_three_steps_positions(name):
    s64, s32 = pair_of_sims(CASES[name])
    s64.run(3)
    s32.run(3)
    keep = s64.get("iam") != sd.DELETED
    r64, r32 = s64.get("r")[keep], s32.get("r")[keep]
    assert np.abs(r32 - r64).max() <= 1e-9
