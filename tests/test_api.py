"""The C ABI used the way a driver uses it: the phase calls with device hooks
between them against sphdem_run_steps (whose hooks are fused), parameter
changes on a live context against the oracle given the same changes, and the
per-kernel timing window."""
import dataclasses

import numpy as np
import pytest

import oracle
import sphdem as sd

pytestmark = pytest.mark.gpu

FIELDS = ("r", "v", "vhat", "f", "fp", "fb", "fdrag", "dens", "h", "press", "porosity")
G = np.array([0.0, 0.0, -9.81])


@pytest.mark.parametrize("precision", [sd.FP64, sd.MIXED])
def test_phase_calls_with_hooks_equal_run_steps(precision):
    """run.cpp's loop spelled out -- sphStep++, start, gravity on sph, middle,
    the sediment3D beforeEnd hooks, end, afterEnd freeze -- ends where
    sphdem_run_steps with the sediment hook set ends, bit for bit (the steps
    take the list and mixed records from the resets' permutation, the phase
    calls pack them apart)."""
    case = sd.sediment_case(nf=12, n_dem=60)
    a, b = sd.Sim(case, precision=precision), sd.Sim(case, precision=precision)
    a.run(3)
    P = case.params
    for _ in range(3):
        g = b.globals()
        g.sphStep += 1
        b.set_globals(g)
        b.start()
        b.hook_body_force(1 << sd.SPH, G)
        b.middle()
        b.hook_sum_drag()
        b.hook_body_force(1 << sd.DEM, G)
        b.hook_dem_floor()
        b.hook_freeze(1 << sd.DEM, P.time_drop)
        b.end()
        b.hook_freeze(1 << sd.DEM, P.time_drop)
    ga, gb = a.globals(), b.globals()
    assert (ga.time, ga.dt, ga.sphStep) == (gb.time, gb.dt, gb.sphStep)
    assert list(ga.custom) == list(gb.custom)
    for nm in FIELDS:
        assert np.array_equal(a.get(nm), b.get(nm)), nm


@pytest.mark.parametrize("precision", [sd.FP64, sd.MIXED])
@pytest.mark.parametrize("hflags", ["var_h", "plain", "const_h"])
def test_phase_calls_equal_run_steps_on_the_closed_box(hflags, precision):
    """The dam-break hook set (gravity on sph before middle) on the closed
    SPH-DEM box, with VAR_H, without it, and with CONST_H: sphdem_run_steps does start() inside
    middle()'s reset and end()'s drift inside end()'s reset; the phase calls
    launch them apart.  Same state bit for bit, a density re-initialisation
    step included."""
    case = sd.dambreak_case(nf=12, with_dem=28, liq_dem_flags=True)
    P = case.params.copy()
    P.reinit_every = 3
    P.flags &= ~(sd.F_VAR_H_CORRECTION | sd.F_CONST_H)
    P.flags |= {"var_h": sd.F_VAR_H_CORRECTION, "plain": 0, "const_h": sd.F_CONST_H}[hflags]
    case = dataclasses.replace(case, params=P)
    a, b = sd.Sim(case, precision=precision), sd.Sim(case, precision=precision)
    a.run(4, hookset=sd.HOOKS_DAMBREAK)
    for _ in range(4):
        g = b.globals()
        g.sphStep += 1
        b.set_globals(g)
        b.start()
        b.hook_body_force(1 << sd.SPH, G)
        b.middle()
        b.end()
    ga, gb = a.globals(), b.globals()
    assert (ga.time, ga.dt, ga.sphStep) == (gb.time, gb.dt, gb.sphStep)
    for nm in FIELDS + ("vhat", "omega", "dddt", "shepSum"):
        assert np.array_equal(a.get(nm), b.get(nm)), nm


def test_phase_calls_equal_run_steps_when_the_reset_deletes():
    """A particle moved above the domain is deleted by middle()'s reset.
    sphdem_run_steps does start() inside that reset, so the particle gets
    start()'s drift and kick there before it is marked deleted; the phase
    calls launch start() first.  Same state bit for bit, the deleted
    particle's fields included."""
    case = sd.sediment_case(nf=12, n_dem=60)
    a, b = sd.Sim(case), sd.Sim(case)
    a.run(2)
    b.run(2)
    for s in (a, b):
        r = s.get("r")
        k = int(np.nonzero(case.iam == sd.SPH)[0][5])
        r[k, 2] = case.params.proc_domain[5] + 0.5 * case.params.h  # leaves in start()'s drift too
        s.set("r", r)
    a.run(2)
    P = case.params
    for _ in range(2):
        g = b.globals()
        g.sphStep += 1
        b.set_globals(g)
        b.start()
        b.hook_body_force(1 << sd.SPH, G)
        b.middle()
        b.hook_sum_drag()
        b.hook_body_force(1 << sd.DEM, G)
        b.hook_dem_floor()
        b.hook_freeze(1 << sd.DEM, P.time_drop)
        b.end()
        b.hook_freeze(1 << sd.DEM, P.time_drop)
    assert a.num_deleted() == b.num_deleted() == 1
    ga, gb = a.globals(), b.globals()
    assert (ga.time, ga.dt, ga.sphStep) == (gb.time, gb.dt, gb.sphStep)
    for nm in FIELDS + ("vhat", "dddt", "iam"):
        assert np.array_equal(a.get(nm), b.get(nm)), nm


def test_parameter_change_mid_run_matches_the_oracle():
    """sphdem_set_params between steps (drag law, stiffness, viscosity) takes
    effect from the next step as in the oracle given the same change."""
    case = sd.sediment_case(nf=12, n_dem=60)
    sim, orc = sd.Sim(case), oracle.Oracle(case)
    sim.run(1)
    orc.run(1)
    P = case.params.copy()
    P.drag_a *= 1.1
    P.dem_k *= 0.9
    P.viscosity *= 1.5
    sim.set_params(P)
    orc.set_params(P)
    sim.run(2)
    orc.run(2)
    assert sim.globals().dt == pytest.approx(orc.globals().dt, rel=1e-12)
    for nm in ("r", "vhat", "dens", "f", "fdrag", "porosity"):
        a, b = sim.get(nm), orc.get(nm)
        err = np.abs(a - b).max() / np.abs(b).max()
        assert err <= 1e-9, (nm, err)
    # the domain of a live context is fixed
    Q = P.copy()
    Q.proc_domain[0] -= 0.1
    with pytest.raises(sd.SphDemError, match="domain"):
        sim.set_params(Q)


def test_kernel_timing_window_counts_launches():
    sim = sd.Sim(sd.sediment_case(nf=12, n_dem=60))
    sim.run(1)
    sim.timing_begin("lists")
    sim.run(3)
    ms, launches = sim.timing_end()
    assert launches == 6 and ms > 0  # two list builds per step (middle and end)
    sim.timing_begin("force")
    sim.run(2)
    ms, launches = sim.timing_end()
    assert launches == 2 and ms > 0
    with pytest.raises(sd.SphDemError, match="unknown kernel"):
        sim.timing_begin("no_such_kernel")


def test_clipped_list_is_reported_by_the_call_that_used_it():
    """A particle with more than the list capacity of SPH neighbours makes the
    call whose pass walked the clipped list fail with SPHDEM_E_CAPACITY (not a
    later reset); the error is reported once and cleared, so the same context
    continues once h is sane again."""
    case = sd.sediment_case(nf=12, n_dem=60)
    print(f"clipped list: sediment nf=12, {case.n} particles", flush=True)
    sim = sd.Sim(case)
    sim.run(1)
    h0 = sim.get("h")
    # h x 2.5: about 15x the ~120 neighbours, far above the 256-entry lists
    sim.set("h", np.where(case.iam == sd.DEM, h0, 2.5 * h0))
    sim.reset()
    with pytest.raises(sd.SphDemError) as e:
        sim.pass_("force")
    assert e.value.code == sd.E_CAPACITY and "neighbours" in str(e.value)
    sim.set("h", h0)
    sim.reset()
    sim.pass_("force")
    # the same through the phase calls: start() recomputes h from the density
    # (driftRest), so the large h is set after it; middle() is the call whose
    # force pass walks the clipped lists
    sim.start()
    sim.set("h", np.where(case.iam == sd.DEM, h0, 2.5 * h0))
    with pytest.raises(sd.SphDemError) as e:
        sim.middle()
    assert e.value.code == sd.E_CAPACITY
    sim.close()
    # and through whole steps under CONST_H (h as set stays): sphdem_run_steps
    # fails in the call that built the clipped lists
    c2 = sd.sediment_case(nf=12, n_dem=60)
    c2.params.flags |= sd.F_CONST_H
    c2.h = np.where(c2.iam == sd.DEM, c2.h, 2.5 * c2.h)
    sim2 = sd.Sim(c2)
    with pytest.raises(sd.SphDemError) as e:
        sim2.run(1)
    assert e.value.code == sd.E_CAPACITY
    sim2.close()
