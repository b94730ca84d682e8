"""Full-state checkpoint (sphdem_checkpoint_save / _restore): a run resumed
from a checkpoint in a fresh context continues bit for bit."""
import numpy as np
import pytest

import sphdem as sd

pytestmark = pytest.mark.gpu

FIELDS = ("r", "v", "vhat", "f", "dens", "h", "omega", "porosity", "shepSum", "fdrag")


@pytest.mark.parametrize("mk", [lambda: sd.sediment_case(nf=12, n_dem=60),
                                lambda: sd.dambreak_case(nf=12)], ids=["sediment", "dambreak"])
def test_resume_is_bit_exact(mk):
    a = sd.Sim(mk())
    a.run(3)
    blob = a.checkpoint()
    a.run(2)
    b = sd.Sim(mk())  # a fresh context with the same parameters
    b.restore(blob)
    b.run(2)
    ga, gb = a.globals(), b.globals()
    assert (ga.time, ga.dt, ga.sphStep) == (gb.time, gb.dt, gb.sphStep)
    for f in FIELDS:
        assert np.array_equal(a.get(f), b.get(f), equal_nan=True), f


def test_restore_rejects_other_parameters():
    a = sd.Sim(sd.sediment_case(nf=12, n_dem=60))
    blob = a.checkpoint()
    other = sd.sediment_case(nf=12, n_dem=60)
    other.params.viscosity *= 2
    b = sd.Sim(other)
    with pytest.raises(sd.SphDemError):
        b.restore(blob)
