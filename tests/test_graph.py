"""Steps replayed from a captured HIP graph (sphdem_run_steps_graph,
sphdem_run_steps_many_graph, sphdem_run_steps_ensemble_graph) and steps
launched with device resets (sphdem_run_steps_many): argument checks without a GPU, and on the GPU
that replayed steps end bit for bit where the same steps run as launches end,
including the steps whose reset needs the host (a deletion stops the replay,
which is rolled back and run as launches) and density reinitialisations."""
import ctypes
import dataclasses

import numpy as np

import pytest

import sphdem as sd

FIELDS = ("r", "v", "vhat", "f", "fp", "fv", "fb", "fdrag", "dens", "h", "press", "porosity",
          "shepSum", "omega", "dddt", "iam")


def same_state(a, b):
    for nm in FIELDS:
        x, y = a.get(nm), b.get(nm)
        assert np.array_equal(x, y, equal_nan=True), nm
    ga, gb = a.globals(), b.globals()
    assert (ga.time, ga.dt, ga.sphStep) == (gb.time, gb.dt, gb.sphStep)
    assert list(ga.custom) == list(gb.custom)


def test_graph_entry_points_refuse_bad_arguments():
    lib = sd.load_library()
    assert lib.sphdem_run_steps_graph(None, 1, 2) == -1
    failed = ctypes.c_int(5)
    assert lib.sphdem_run_steps_many_graph(None, 0, 3, 2, 0, ctypes.byref(failed)) == 0
    assert failed.value == -1
    fake = ctypes.c_void_p(0x1000)
    twice = (ctypes.c_void_p * 2)(fake.value, fake.value)
    assert lib.sphdem_run_steps_many_graph(twice, 2, 1, 2, 0, None) == -1
    one = (ctypes.c_void_p * 1)(fake.value)
    assert lib.sphdem_run_steps_many_graph(one, 1, -1, 2, 0, None) == -1
    assert lib.sphdem_run_steps_many_graph(one, 1, 1, 7, 0, None) == -1
    for fn in (lib.sphdem_run_steps_many_graph, lib.sphdem_run_steps_ensemble_graph):
        assert fn(twice, 2, 1, 2, 0, None) == -1
        assert fn(one, 1, 1, 7, 0, None) == -1
    with pytest.raises(ValueError):
        sd.run_many([object()], 1, hookset=2, graph="boxes")


CASES = {
    "sediment": lambda: sd.sediment_case(nf=12, n_dem=60),
    "closedbox": lambda: sd.dambreak_case(nf=12, with_dem=28, liq_dem_flags=True),
}


@pytest.mark.gpu
@pytest.mark.parametrize("precision", [sd.FP64, sd.MIXED])
@pytest.mark.parametrize("name", sorted(CASES))
def test_graph_steps_equal_launched_steps(name, precision):
    """A host step, a capture, then replays in two calls: the same state as
    the same steps run as launches, bit for bit."""
    case = CASES[name]()
    a, b = sd.Sim(case, precision=precision), sd.Sim(case, precision=precision)
    a.run(6)
    b.run_graph(4)
    b.run_graph(2)
    same_state(a, b)
    a.close()
    b.close()


@pytest.mark.gpu
def test_a_deletion_stops_the_replay_and_the_step_runs_as_launches():
    """A particle moved out of the domain between calls: the next replayed
    step's reset deletes it, which only the host can handle, so the replay is
    stopped, the state before it restored and the step run as launches; the
    run then captures again.  Same state as launches throughout."""
    case = sd.sediment_case(nf=12, n_dem=60)
    a, b = sd.Sim(case), sd.Sim(case)
    a.run(3)
    b.run_graph(3)
    same_state(a, b)
    for s in (a, b):
        r = s.get("r")
        k = int(np.nonzero(case.iam == sd.SPH)[0][7])
        r[k, 2] = case.params.proc_domain[5] + 1.0  # above the domain: deleted
        s.set("r", r)
    a.run(4)
    b.run_graph(4)
    assert a.num_deleted() == b.num_deleted() == 1
    same_state(a, b)
    a.close()
    b.close()


@pytest.mark.gpu
def test_reinit_steps_run_as_launches_between_replays():
    case = sd.dambreak_case(nf=12, with_dem=28, liq_dem_flags=True)
    P = case.params.copy()
    P.reinit_every = 3
    case2 = dataclasses.replace(case, params=P)
    a, b = sd.Sim(case2), sd.Sim(case2)
    a.run(7)
    b.run_graph(7)
    same_state(a, b)
    a.close()
    b.close()


@pytest.mark.gpu
@pytest.mark.parametrize("graph", [True, "ensemble"])
def test_many_graph_equals_many(graph):
    """Replays of each box's own graph, or of one graph over all boxes (every
    box replaying the same count; one box captured again after a tile-shape
    change replays a step fewer, and the call falls back to the boxes' own
    graphs), end where the same steps as launches end."""
    cases = sd.ensemble_cases(nbox=5, seed=91)
    a = [sd.Sim(c, precision=sd.MIXED) for c in cases]
    b = [sd.Sim(c, precision=sd.MIXED) for c in cases]
    sd.run_many(a, 4, workers=2)
    sd.run_many(b, 4, workers=2, graph=graph)
    sd.run_many(b, 1, workers=3, graph=graph)
    sd.run_many(a, 1, workers=3)
    for s in (a[3], b[3]):
        s.set_tile_shape(64, 128)  # box 3's capture no longer matches
    sd.run_many(b, 3, workers=2, graph=graph)
    sd.run_many(a, 3, workers=2)
    sd.run_many(b, 2, workers=2, graph=graph)
    sd.run_many(a, 2, workers=2)
    for x, y in zip(a, b):
        same_state(x, y)
        x.close()
        y.close()


@pytest.mark.gpu
@pytest.mark.parametrize("graph", [False, True, "ensemble"])
def test_an_ensemble_box_whose_reset_needs_the_host_recovers(graph):
    """sphdem_run_steps_many launches each box's steps with device resets and
    reads nothing back until the end (sphdem_run_steps_many_graph replays each
    box's captured step, sphdem_run_steps_ensemble_graph one graph holding
    them all): a box with a particle moved out
    of its domain is stopped at that step's reset, rolled back, and finishes
    as launches.  Every box ends as it does stepped alone."""
    cases = sd.ensemble_cases(nbox=4, seed=93)
    a = [sd.Sim(c, precision=sd.MIXED) for c in cases]
    b = [sd.Sim(c, precision=sd.MIXED) for c in cases]
    sd.run_many(b, 2, workers=2, graph=graph)
    for s in a:
        s.run(2)
    for s, c in ((a[2], cases[2]), (b[2], cases[2])):
        r = s.get("r")
        k = int(np.nonzero(c.iam == sd.SPH)[0][3])
        r[k, 2] = c.params.proc_domain[5] + 1.0
        s.set("r", r)
    sd.run_many(b, 3, workers=2, graph=graph)
    for s in a:
        s.run(3)
    assert a[2].num_deleted() == b[2].num_deleted() == 1
    for x, y in zip(a, b):
        same_state(x, y)
        x.close()
        y.close()


@pytest.mark.gpu
@pytest.mark.skip(reason="batched ensemble launches (each box's step recorded once, launch i "
                         "covering every box's i-th launch) faulted the GPU in their first "
                         "run; the cause is not found and the path is not in this build")
@pytest.mark.parametrize("precision", [sd.FP64, sd.MIXED])
def test_batched_ensemble_steps_equal_boxes_alone(precision):
    """sphdem_run_steps_many with every box ready and the same steps left
    records each box's step once and launches the steps batched over all
    boxes (one box with another walker block: its walker launches form a
    group of their own).  Same state as each box stepped alone, bit for bit,
    over two calls that reuse the recordings."""
    cases = sd.ensemble_cases(nbox=6, seed=95)
    a = [sd.Sim(c, precision=precision) for c in cases]
    b = [sd.Sim(c, precision=precision) for c in cases]
    for s in (a[1], b[1]):
        s.set_tile_shape(0, 128)
    sd.run_many(b, 1, workers=3)  # each box's first step, as launches
    sd.run_many(b, 3, workers=3)
    assert all(s.last_steps_batched() for s in b)
    sd.run_many(b, 2, workers=2)
    assert all(s.last_steps_batched() for s in b)
    for s in a:
        s.run(6)
    for x, y in zip(a, b):
        same_state(x, y)
        x.close()
        y.close()


@pytest.mark.gpu
def test_an_ensemble_with_a_segmented_list_build_still_equals_boxes_alone():
    """A box whose list build is segmented launches other kernels than the
    rest; batched or not, every box ends where it ends stepped alone."""
    cases = sd.ensemble_cases(nbox=4, seed=97)
    a = [sd.Sim(c, precision=sd.MIXED) for c in cases]
    b = [sd.Sim(c, precision=sd.MIXED) for c in cases]
    for s in (a[2], b[2]):
        s.set_tile_shape(32, 256)
    sd.run_many(b, 4, workers=2)
    sd.run_many(b, 2, workers=2)
    for s in a:
        s.run(6)
    for x, y in zip(a, b):
        same_state(x, y)
        x.close()
        y.close()
