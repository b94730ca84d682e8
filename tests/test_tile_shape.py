"""Tile shapes (particles per list-build segment or 0 for the per-target
search, threads per walker block) and their autotuner: results do not depend
on the shape, bit for bit, and tuning leaves the simulation state as it was."""
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
@pytest.mark.parametrize("shape", [(64, 256), (32, 64), (48, 128), (8, 256), (0, 128)])
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
    assert seg in (0, 64, 48, 32) and blk in (256, 128, 64) and ms > 0
    assert a.tile_shape() == (seg, blk)
    a.run(2)
    b.run(2)
    for nm in FIELDS:
        assert np.array_equal(a.get(nm), b.get(nm)), nm
    assert a.globals().sphStep == b.globals().sphStep == 3
    a.close()
    b.close()


def _pairs_for(case, seg, edit=None):
    sim = sd.Sim(case)
    sim.set_tile_shape(seg, 256)
    if edit:
        edit(sim)
    sim.reset()
    p = sim.pairs(sd.PAIRS_SPH)
    sim.close()
    return p


@pytest.mark.gpu
@pytest.mark.parametrize("spread", [0.0, 0.3])
def test_per_target_search_equals_the_segmented_search(spread):
    """The per-target search decides most candidates on 16-bit quantised
    positions and the rest by the FP64 test: its lists equal the segmented
    search's (FP64 test on every candidate) entry by entry, with h uniform and
    with h spread over [1 - spread, 1] of its value (many pairs decided by the
    larger h of the two), on a periodic case (ghosts) and a closed box."""
    rng = np.random.default_rng(11)
    for case in (sd.sediment_case(nf=16, n_dem=64, seed=3),
                 sd.dambreak_case(nf=16, dx=0.32 / 16, seed=4, with_dem=100, liq_dem_flags=True)):
        h = None

        def edit(sim):
            nonlocal h
            if h is None:
                h0 = sim.get("h")
                h = h0 * (1.0 - spread * rng.random(h0.shape))
            sim.set("h", h)

        a = _pairs_for(case, 64, edit if spread else None)
        b = _pairs_for(case, 0, edit if spread else None)
        assert a.shape == b.shape and np.array_equal(a, b)


@pytest.mark.gpu
def test_per_target_search_with_a_zero_h_falls_back_to_the_fp64_test():
    """A non-positive h (or a non-finite position) anywhere makes the records
    unusable as bounds: every candidate goes to the FP64 test, and the lists
    still equal the segmented search's."""
    case = sd.sediment_case(nf=12, n_dem=60)

    def edit(sim):
        h = sim.get("h")
        h[5] = 0.0
        sim.set("h", h)

    a = _pairs_for(case, 64, edit)
    b = _pairs_for(case, 0, edit)
    assert np.array_equal(a, b)


@pytest.mark.gpu
def test_switching_to_a_segmented_search_after_a_per_target_reset():
    """Segments are built only for a segmented search: a switch from the
    per-target search between a reset and the list build builds them then."""
    case = sd.sediment_case(nf=12, n_dem=60)
    ref = _pairs_for(case, 64)
    sim = sd.Sim(case)
    sim.set_tile_shape(0, 256)
    sim.reset()
    sim.set_tile_shape(48, 256)
    assert np.array_equal(sim.pairs(sd.PAIRS_SPH), ref)
    sim.run(1)
    sim.close()
