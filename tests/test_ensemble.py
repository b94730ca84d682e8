"""Ensembles of independent boxes (sphdem_run_steps_many): the box generator,
argument checks without a GPU, and on the GPU that boxes advanced together
end bit for bit where each box run alone ends."""
import ctypes
import math

import numpy as np
import pytest

import sphdem as sd


def test_ensemble_boxes_follow_the_draws():
    cases = sd.ensemble_cases(nbox=12, seed=77)
    assert len(cases) == 12
    sizes = []
    for c in cases:
        nf = int((c.iam == sd.SPH).sum())
        nd = int((c.iam == sd.DEM).sum())
        m, nz = c.info["nf"], c.info["nz"]
        assert nf == m * m * nz
        assert nd == nf // 16
        # m^2 nz rounds the drawn count: within a few percent of [4096, 65536]
        assert 0.9 * 4096 <= nf <= 1.1 * 65536
        for k in ("scale_drag", "scale_k", "scale_gamma"):
            assert 0.8 <= c.info[k] <= 1.2
        P = c.params
        assert P.drag_a == pytest.approx(3.7 * c.info["scale_drag"], rel=1e-15)
        assert P.drag_b == pytest.approx(0.65 * c.info["scale_drag"], rel=1e-15)
        # DEM_TIMESTEP = t_c / 50 of the scaled spring-dashpot (demConstants.h:6)
        m_red = 0.5 * P.dem_vol * P.dem_dens
        t_c = sd.PI_REF / math.sqrt(P.dem_k / m_red - (0.5 * P.dem_gamma / m_red) ** 2)
        assert P.dem_timestep == pytest.approx(t_c / 50.0, rel=1e-14)
        # the box is nz dx high, DEM in its upper half
        lz = P.rmax[2]
        assert lz == pytest.approx(nz / m, rel=1e-15)
        dem = c.r[c.iam == sd.DEM]
        assert dem[:, 2].min() >= 0.5 * lz and dem[:, 2].max() <= lz
        sizes.append(nf)
    assert len(set(sizes)) > 1  # uneven sizes
    # reproducible
    again = sd.ensemble_cases(nbox=12, seed=77)
    assert all(np.array_equal(a.r, b.r) for a, b in zip(cases, again))


def test_run_steps_many_rejects_bad_arguments():
    lib = sd.load_library()
    failed = ctypes.c_int(5)
    assert lib.sphdem_run_steps_many(None, 0, 3, 2, 0, ctypes.byref(failed)) == 0
    assert failed.value == -1
    assert lib.sphdem_run_steps_many(None, 2, 3, 2, 0, None) == -1
    fake = ctypes.c_void_p(0x1000)
    twice = (ctypes.c_void_p * 2)(fake.value, fake.value)
    assert lib.sphdem_run_steps_many(twice, 2, 1, 2, 0, None) == -1  # repeated context
    null = (ctypes.c_void_p * 2)(fake.value, None)
    assert lib.sphdem_run_steps_many(null, 2, 1, 2, 0, None) == -1
    one = (ctypes.c_void_p * 1)(fake.value)
    assert lib.sphdem_run_steps_many(one, 1, -1, 2, 0, None) == -1
    assert lib.sphdem_run_steps_many(one, 1, 1, 2, -3, None) == -1


@pytest.mark.gpu
@pytest.mark.parametrize("precision", [sd.FP64, sd.MIXED])
def test_boxes_together_equal_boxes_alone(precision):
    cases = sd.ensemble_cases(nbox=5, seed=31, nmin=512, nmax=4096)
    together = [sd.Sim(c, precision=precision) for c in cases]
    sd.run_many(together, 3, workers=3)
    for c, s in zip(cases, together):
        alone = sd.Sim(c, precision=precision)
        alone.run(3)
        ga, gt = alone.globals(), s.globals()
        assert (gt.time, gt.dt, gt.sphStep) == (ga.time, ga.dt, ga.sphStep)
        assert list(gt.custom) == list(ga.custom)
        for nm in ("r", "v", "vhat", "dens", "h", "f", "porosity", "fdrag", "shepSum"):
            assert np.array_equal(s.get(nm), alone.get(nm)), (c.info["box"], nm)
        alone.close()
    for s in together:
        s.close()


@pytest.mark.gpu
def test_a_failing_box_is_reported():
    cases = sd.ensemble_cases(nbox=3, seed=32, nmin=512, nmax=1024)
    sims = [sd.Sim(c) for c in cases]
    with pytest.raises(sd.SphDemError, match="box"):
        sd.run_many(sims, 1, hookset=7, workers=2)  # no such hook set
    for s in sims:
        s.close()


@pytest.mark.gpu
def test_ensemble_boxes_match_the_oracle():
    """Boxes of the ensemble generator (their own drag, stiffness and damping
    draws) advanced together by sphdem_run_steps_many in the FP64 form, each
    compared with the oracle run of that box: pair sets after the run
    bit-exact, fields within the whole-step bound of the parity tests."""
    import oracle

    cases = sd.ensemble_cases(nbox=3, seed=41, nmin=1000, nmax=3000)
    sims = [sd.Sim(c) for c in cases]
    sd.run_many(sims, 2, workers=3)
    for c, s in zip(cases, sims):
        print(f"ensemble box {c.info['box']}: {c.counts()}", flush=True)
        orc = oracle.Oracle(c)
        orc.run(2)
        alive = orc.get("iam") != -1
        for nm in ("r", "vhat", "dens", "h", "f", "porosity", "fdrag", "shepSum"):
            a, b = s.get(nm)[alive], orc.get(nm)[alive]
            err = float(np.abs(a - b).max() / max(np.abs(b).max(), 1e-300))
            assert err <= 1e-10, (c.info["box"], nm, err)
        # the searches of the evolved state (a reset on both first: the cached
        # lists predate end()'s driftRest)
        s.reset()
        orc.reset()
        for kind in (sd.PAIRS_SPH, sd.PAIRS_SPH2DEM, sd.PAIRS_DEM2SPH):
            g = set(map(tuple, s.pairs(kind).tolist()))
            o = set(map(tuple, orc.pairs(kind).tolist()))
            assert g == o, (c.info["box"], kind)
        orc.close()
        s.close()
