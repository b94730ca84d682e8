"""Writes the golden fixtures of the timestep loop: the CPU oracle's state
after two steps of the two small test cases, from the seeded initial
conditions of sphdem.py.  The reference ships no golden vectors for this
path and does not build here, so these pin the oracle against itself (any
change to it or to the case generators shows up in tests/test_golden.py)
and give the GPU path a fixed target.

    python tests/golden/make_golden.py      # rewrites tests/golden/*.npz
"""
import hashlib
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, os.path.join(ROOT, "sph-dem_amd"))
sys.path.insert(0, os.path.join(ROOT, "tests"))

import sphdem as sd  # noqa: E402

CASES = {
    "dambreak12": lambda: sd.dambreak_case(nf=12),
    "sediment12": lambda: sd.sediment_case(nf=12, n_dem=60),
}
STEPS = 2
FIELDS = ("r", "v", "dens", "h", "f", "porosity")


def ic_digest(c: sd.Case) -> str:
    h = hashlib.sha256()
    for a in (c.r, c.v, c.vhat, c.mass, c.h, c.dens, c.iam):
        h.update(np.ascontiguousarray(a).tobytes())
    h.update(bytes(c.params))
    return h.hexdigest()


def main():
    import oracle

    for name, mk in CASES.items():
        c = mk()
        o = oracle.Oracle(c)
        o.run(STEPS)
        g = o.globals()
        out = {f: o.get(f) for f in FIELDS}
        out["time"] = np.float64(g.time)
        out["dt"] = np.float64(g.dt)
        out["steps"] = np.int64(STEPS)
        out["ic_sha256"] = np.bytes_(ic_digest(c))
        np.savez_compressed(os.path.join(HERE, f"{name}.npz"), **out)
        o.close()
        print(name, c.n, "particles")


if __name__ == "__main__":
    main()
