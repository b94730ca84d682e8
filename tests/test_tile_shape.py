"""Tile shapes (particles per list-build segment, threads per walker block)
and their autotuner: results do not depend on the shape, bit for bit, and
tuning leaves the simulation state as it was."""
import ctypes

import numpy as np
import pytest

import sphdem as sd

FIELDS = ("r", "v", "vhat", "f", "dens", "h", "press", "porosity", "fdrag", "shepSum", "omega")


def test_tile_shape_entry_points_refuse_a_null_context():
    lib = sd.load_library()
    s, b = ctypes.c_int(), ctypes.c_int()
    assert lib.sphdem_set_tile_shape(None, 64, 256) == -1
    assert lib.sphdem_get_tile_shape(None, ctypes.byref(s), ctypes.byref(b)) == -1
    assert lib.sphdem_tune_tile_shape(None, 2, 2, ctypes.byref(s), ctypes.byref(b), None) == -1


@pytest.mark.gpu
@pytest.mark.parametrize("shape", [(32, 64), (48, 128), (8, 256)])
def test_results_do_not_depend_on_the_tile_shape(shape):
    case = sd.sediment_case(nf=12, n_dem=60)
    ref = sd.Sim(case)
    ref.run(3)
    sim = sd.Sim(case)
    sim.set_tile_shape(*shape)
    assert sim.tile_shape() == shape
    sim.run(3)
    for nm in FIELDS:
        assert np.array_equal(sim.get(nm), ref.get(nm)), (shape, nm)
    assert np.array_equal(sim.pairs(sd.PAIRS_SPH), ref.pairs(sd.PAIRS_SPH))
    assert (sim.globals().dt, sim.globals().time) == (ref.globals().dt, ref.globals().time)
    with pytest.raises(sd.SphDemError):
        sim.set_tile_shape(65, 256)
    with pytest.raises(sd.SphDemError):
        sim.set_tile_shape(64, 96)
    ref.close()
    sim.close()


@pytest.mark.gpu
def test_tuning_leaves_the_state_as_it_was():
    case = sd.sediment_case(nf=12, n_dem=60)
    a, b = sd.Sim(case), sd.Sim(case)
    a.run(1)
    b.run(1)
    seg, blk, ms = a.tune_tile_shape(steps=1)
    assert seg in (64, 48, 32) and blk in (256, 128, 64) and ms > 0
    assert a.tile_shape() == (seg, blk)
    a.run(2)
    b.run(2)
    for nm in FIELDS:
        assert np.array_equal(a.get(nm), b.get(nm)), nm
    assert a.globals().sphStep == b.globals().sphStep == 3
    a.close()
    b.close()
