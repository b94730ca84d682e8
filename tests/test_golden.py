"""The committed golden fixtures (tests/golden/*.npz, written by
tests/golden/make_golden.py): the oracle's state after two steps of the two
small test cases.  On the CPU the oracle must reproduce them bit for bit and
the case generators must still produce the same inputs; on the GPU the
library must reach them within the multi-step tolerance."""
import os
import sys

import numpy as np
import pytest

import oracle
import sphdem as sd

HERE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
sys.path.insert(0, HERE)
import make_golden  # noqa: E402

NAMES = sorted(make_golden.CASES)


def load(name):
    return np.load(os.path.join(HERE, f"{name}.npz"), allow_pickle=False)


@pytest.mark.parametrize("name", NAMES)
def test_inputs_unchanged(name):
    g = load(name)
    c = make_golden.CASES[name]()
    assert make_golden.ic_digest(c) == g["ic_sha256"].item().decode()


@pytest.mark.parametrize("name", NAMES)
def test_oracle_reproduces_fixture(name):
    g = load(name)
    c = make_golden.CASES[name]()
    o = oracle.Oracle(c)
    o.run(int(g["steps"]))
    gl = o.globals()
    assert gl.time == g["time"] and gl.dt == g["dt"]
    for f in make_golden.FIELDS:
        assert np.array_equal(o.get(f), g[f], equal_nan=True), f
    o.close()


@pytest.mark.gpu
@pytest.mark.parametrize("name", NAMES)
def test_gpu_reaches_fixture(name):
    """Two steps on the GPU against the fixture: 1e-10 normwise per field."""
    g = load(name)
    c = make_golden.CASES[name]()
    sim = sd.Sim(c)
    sim.run(int(g["steps"]))
    gl = sim.globals()
    assert gl.time == pytest.approx(float(g["time"]), rel=1e-12)
    assert gl.dt == pytest.approx(float(g["dt"]), rel=1e-12)
    for f in make_golden.FIELDS:
        a, b = sim.get(f), g[f]
        keep = np.isfinite(b).all(axis=-1) if b.ndim > 1 else np.isfinite(b)
        scale = np.abs(b[keep]).max()
        err = np.abs(a[keep] - b[keep]).max() / scale
        assert err <= 1e-10, f"{f}: {err}"
    sim.close()
