"""ctypes driver of the CPU oracle (oracle/liboracle.so).  Test infrastructure:
the oracle is the checker the GPU library is compared with."""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

import sphdem as sd

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE_DIR = os.path.join(ROOT, "oracle")
ORACLE_SO = os.path.join(ORACLE_DIR, "liboracle.so")

_lib = None


def load():
    global _lib
    if _lib is not None:
        return _lib
    src = os.path.join(ORACLE_DIR, "sphdem_oracle.c")
    if (not os.path.exists(ORACLE_SO)) or os.path.getmtime(ORACLE_SO) < os.path.getmtime(src):
        subprocess.run(["make", "-s", "-C", ORACLE_DIR], check=True)
    lib = ctypes.CDLL(ORACLE_SO)
    vp = ctypes.c_void_p
    dp = ctypes.POINTER(ctypes.c_double)
    ip = ctypes.POINTER(ctypes.c_int32)
    lib.orc_create.restype = vp
    lib.orc_create.argtypes = [ctypes.POINTER(sd.Params), ctypes.c_int, dp, dp, dp, dp, dp, dp, ip]
    lib.orc_destroy.argtypes = [vp]
    lib.orc_set_params.argtypes = [vp, ctypes.POINTER(sd.Params)]
    for nm in ("orc_num_particles", "orc_num_ghosts", "orc_reset", "orc_start", "orc_middle",
               "orc_end"):
        getattr(lib, nm).argtypes = [vp]
        getattr(lib, nm).restype = ctypes.c_int
    lib.orc_num_deleted.argtypes = [vp]
    lib.orc_num_deleted.restype = ctypes.c_long
    lib.orc_search_radius.argtypes = [vp]
    lib.orc_search_radius.restype = ctypes.c_double
    lib.orc_set_globals.argtypes = [vp, ctypes.POINTER(sd.Globals)]
    lib.orc_get_globals.argtypes = [vp, ctypes.POINTER(sd.Globals)]
    lib.orc_get_field.argtypes = [vp, ctypes.c_int, vp]
    lib.orc_set_field.argtypes = [vp, ctypes.c_int, vp]
    lib.orc_refresh_ghosts.argtypes = [vp]
    for nm in ("omega", "eos", "force", "dddt", "density", "porosity", "fpi", "dem_contacts"):
        getattr(lib, "orc_pass_" + nm).argtypes = [vp]
    lib.orc_new_dt.argtypes = [vp]
    lib.orc_new_dt.restype = ctypes.c_double
    lib.orc_set_new_dt.argtypes = [vp, ctypes.c_double]
    lib.orc_hook_body_force.argtypes = [vp, ctypes.c_uint32, dp]
    lib.orc_hook_dem_floor.argtypes = [vp]
    lib.orc_hook_freeze.argtypes = [vp, ctypes.c_uint32, ctypes.c_double]
    lib.orc_hook_sum_drag.argtypes = [vp]
    lib.orc_run_steps.argtypes = [vp, ctypes.c_int, ctypes.c_int]
    lib.orc_run_steps.restype = ctypes.c_int
    lib.orc_cells.argtypes = [vp, ip]
    lib.orc_pairs.argtypes = [vp, ctypes.c_int, ip, ip, ip, ctypes.c_long]
    lib.orc_pairs.restype = ctypes.c_long
    for nm in ("W", "F", "K", "pdWh"):
        getattr(lib, "orc_kernel_" + nm).argtypes = [vp, ctypes.c_double, ctypes.c_double]
        getattr(lib, "orc_kernel_" + nm).restype = ctypes.c_double
    lib.orc_particle_drag.argtypes = [vp, ctypes.c_double, dp, dp, ctypes.c_double, dp]
    lib.orc_output_vars.argtypes = [vp, ctypes.POINTER(sd.Output)]
    lib.orc_set_threads.argtypes = [ctypes.c_int]
    lib.orc_set_threads.restype = ctypes.c_int
    _lib = lib
    return lib


def set_threads(n: int) -> int:
    """Host threads of the oracle's parallel gathers (0: leave as is); returns
    the number in use.  Results do not depend on it."""
    return load().orc_set_threads(n)


def _dp(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


class Oracle:
    """The CPU restatement holding one case (particles indexed by original index)."""

    def __init__(self, case: sd.Case, params: sd.Params | None = None):
        self.lib = load()
        self.case = case
        self.params = (params or case.params).copy()
        self._keep = [np.ascontiguousarray(x, dtype=np.float64)
                      for x in (case.r, case.v, case.vhat, case.mass, case.h, case.dens)]
        iam = np.ascontiguousarray(case.iam, dtype=np.int32)
        self._keep.append(iam)
        self.h = self.lib.orc_create(ctypes.byref(self.params), case.n,
                                     *[_dp(x) for x in self._keep[:6]],
                                     iam.ctypes.data_as(ctypes.POINTER(ctypes.c_int32)))
        if not self.h:
            raise RuntimeError("oracle construction failed")
        g = sd.Globals()
        g.maxdt = self.params.maxdt
        self.set_globals(g)

    def close(self):
        if self.h:
            self.lib.orc_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def get(self, name: str) -> np.ndarray:
        n = self.case.n
        if name in sd.INT_FIELDS:
            out = np.full(n, -1, dtype=np.int32)
        elif name in sd.VECTOR_FIELDS:
            out = np.full((n, 3), np.nan)
        else:
            out = np.full(n, np.nan)
        assert self.lib.orc_get_field(self.h, sd.FIELD[name], out.ctypes.data) == 0
        return out

    def set(self, name: str, value):
        dt = np.int32 if name in sd.INT_FIELDS else np.float64
        a = np.ascontiguousarray(value, dtype=dt)
        assert self.lib.orc_set_field(self.h, sd.FIELD[name], a.ctypes.data) == 0

    def refresh_ghosts(self):
        self.lib.orc_refresh_ghosts(self.h)

    def globals(self) -> sd.Globals:
        g = sd.Globals()
        self.lib.orc_get_globals(self.h, ctypes.byref(g))
        return g

    def set_globals(self, g: sd.Globals):
        self.lib.orc_set_globals(self.h, ctypes.byref(g))

    def set_params(self, P: sd.Params):
        self.params = P.copy()
        self.lib.orc_set_params(self.h, ctypes.byref(self.params))

    def reset(self):
        assert self.lib.orc_reset(self.h) == 0

    def start(self):
        assert self.lib.orc_start(self.h) == 0

    def middle(self):
        assert self.lib.orc_middle(self.h) == 0

    def end(self):
        assert self.lib.orc_end(self.h) == 0

    def run(self, nsteps: int, hookset: int | None = None):
        hs = self.case.hookset if hookset is None else hookset
        assert self.lib.orc_run_steps(self.h, nsteps, hs) == 0

    def pass_(self, name: str):
        getattr(self.lib, "orc_pass_" + name)(self.h)

    def new_dt(self) -> float:
        return self.lib.orc_new_dt(self.h)

    def set_new_dt(self, v: float):
        self.lib.orc_set_new_dt(self.h, v)

    def hook_body_force(self, mask, accel):
        a = np.ascontiguousarray(accel, dtype=np.float64)
        self.lib.orc_hook_body_force(self.h, mask, _dp(a))

    def hook_dem_floor(self):
        self.lib.orc_hook_dem_floor(self.h)

    def hook_freeze(self, mask, t):
        self.lib.orc_hook_freeze(self.h, mask, t)

    def hook_sum_drag(self):
        self.lib.orc_hook_sum_drag(self.h)

    def cells(self) -> np.ndarray:
        out = np.full((self.case.n, 3), -1, dtype=np.int32)
        self.lib.orc_cells(self.h, out.ctypes.data_as(ctypes.POINTER(ctypes.c_int32)))
        return out

    def pairs(self, kind: int) -> np.ndarray:
        n = self.lib.orc_pairs(self.h, kind, None, None, None, 0)
        a = np.empty(max(n, 1), dtype=np.int32)
        b = np.empty(max(n, 1), dtype=np.int32)
        im = np.empty(max(n, 1), dtype=np.int32)
        ip = ctypes.POINTER(ctypes.c_int32)
        self.lib.orc_pairs(self.h, kind, a.ctypes.data_as(ip), b.ctypes.data_as(ip),
                           im.ctypes.data_as(ip), n)
        return np.stack([a[:n], b[:n], im[:n]], axis=1)

    def output_vars(self) -> dict:
        """calcOutputVars diagnostics as a dict (the fields of sphdem_output)."""
        o = sd.Output()
        self.lib.orc_output_vars(self.h, ctypes.byref(o))
        d = {k: getattr(o, k) for k, _ in sd.Output._fields_}
        d["linMom"] = list(o.linMom)
        return d

    def num_ghosts(self) -> int:
        return self.lib.orc_num_ghosts(self.h)

    def num_particles(self) -> int:
        return self.lib.orc_num_particles(self.h)

    def num_deleted(self) -> int:
        return self.lib.orc_num_deleted(self.h)

    def search_radius(self) -> float:
        return self.lib.orc_search_radius(self.h)

    # kernel family for closed-form tests
    def W(self, q, h):
        return self.lib.orc_kernel_W(self.h, q, h)

    def F(self, q, h):
        return self.lib.orc_kernel_F(self.h, q, h)

    def K(self, q, h):
        return self.lib.orc_kernel_K(self.h, q, h)

    def pdWh(self, q, h):
        return self.lib.orc_kernel_pdWh(self.h, q, h)

    def particle_drag(self, porosity, vhat, fvel, mass):
        out = np.zeros(3)
        v = np.ascontiguousarray(vhat, dtype=np.float64)
        fv = np.ascontiguousarray(fvel, dtype=np.float64)
        self.lib.orc_particle_drag(self.h, porosity, _dp(v), _dp(fv), mass, _dp(out))
        return out
