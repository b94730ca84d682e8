"""calcOmega summed inside the step's force walk (middle) and dddt walk (end).

Under VAR_H_CORRECTION the step computes omega in the walk that reads it
instead of in a pass of its own.  middle() is replayed here with the
per-pass entry points from a checkpoint taken right before it: a reset (the
same sort and the same lists), the coupling pass, omega, EOS and force.
Omega must equal the separate pass bit for bit (the same sum in list order).
f, fp and fv differ only in rounding, because the fused walk scales its sums
by 1/omega once instead of each term: 1e-12 normwise in the FP64 form, and
1e-6 in the mixed form, where the separate pass applies an FP32 1/omega per
term."""
import numpy as np
import pytest

import sphdem as sd

pytestmark = pytest.mark.gpu

G = np.array([0.0, 0.0, -9.81])
TOL = {sd.FP64: 1e-12, sd.MIXED: 1e-6}


def normwise(a, b):
    scale = np.abs(b).max()
    return float(np.abs(a - b).max() / scale) if scale > 0 else float(np.abs(a).max())


def sediment():
    return sd.sediment_case(nf=12, n_dem=60)


def dam_artificial_varh():
    c = sd.dambreak_case(nf=12)
    c.params.flags |= sd.F_VAR_H_CORRECTION
    return c


def dam_monaghan_varh():
    c = sd.dambreak_case(nf=12)
    c.params.flags = ((c.params.flags & ~(sd.F_VISC_MORRIS | sd.F_VISC_ARTIFICIAL))
                      | sd.F_VISC_MONAGHAN | sd.F_VAR_H_CORRECTION)
    return c


CASES = {"sediment": sediment, "dam_artificial_varh": dam_artificial_varh,
         "dam_monaghan_varh": dam_monaghan_varh}
FORMS = {"fp64": sd.FP64, "mixed": sd.MIXED}


@pytest.mark.parametrize("form", list(FORMS))
@pytest.mark.parametrize("name", list(CASES))
def test_middle_omega_and_force_equal_the_separate_passes(name, form):
    case = CASES[name]()
    prec = FORMS[form]
    assert case.params.flags & sd.F_VAR_H_CORRECTION
    print(f"fused omega: {name}, {form}", flush=True)
    a, b = sd.Sim(case, precision=prec), sd.Sim(case, precision=prec)
    a.run(2)
    g = a.globals()
    g.sphStep += 1
    a.set_globals(g)
    a.start()
    a.hook_body_force(1 << sd.SPH, G)
    blob = a.checkpoint()
    a.middle()  # omega inside the force walk

    b.restore(blob)
    b.reset()
    if case.params.flags & sd.F_LIQ_DEM:
        b.pass_("porosity")
    b.pass_("omega")
    b.pass_("eos")
    b.pass_("force")

    sph = (case.iam == sd.SPH) | (case.iam == sd.SPH_BOUNDARY)
    om_a, om_b = a.get("omega")[sph], b.get("omega")[sph]
    assert np.all(np.isfinite(om_a)) and np.abs(om_a - 1).max() > 0
    assert np.array_equal(om_a, om_b), f"omega differs by {np.abs(om_a - om_b).max():.3e}"
    for nm in ("f", "fp", "fv"):
        x, y = a.get(nm)[sph], b.get(nm)[sph]
        err = normwise(x, y)
        assert err <= TOL[prec], f"{nm}: {err:.3e}"
    assert np.abs(b.get("fp")[sph]).max() > 0
    a.close()
    b.close()
