"""Adjoint of the SPH-DEM timestep loop: gradients of a loss on the DEM
positions after a rollout with respect to the Di Felice drag-closure
coefficients, the contact stiffness and damping, the fluid viscosity and the
initial particle velocities.

The reference has no derivative code.  This module restates its FP64 step for
the sediment3D case -- CsphIncompress::start/middle/end (sphIncompress.cpp:
3-251) with the sediment3D hooks (sediment3D/customSim.cpp:5-113) -- as a
differentiable function of the state, in original particle order, and takes
its gradient by reverse mode:

* Every pair pass -- the SPH force, omega, dddt, the fluid-side coupling sums
  (porosity, Shepard sums, DEM->SPH drag), the fluid-particle drag (calcFPI)
  and the DEM contacts -- runs as a `torch.autograd.Function` whose forward
  and backward are the hand-written HIP kernels of include/sphdem_adjoint.h
  (`HipPasses`).
* The per-particle updates (integrator, EOS, h from density, porosity
  clamp, the dt rule, the hooks) are torch expressions on the device,
  differentiated by torch.
* Neighbour sets are discrete: at every reset they come from the library's own
  searches (`LibraryPairs`: a helper context given the model's positions and
  h), which reproduces the reference's pair sets bit for bit, including the
  cell rule of the SPH->DEM coupling search.  They carry no gradient, as in
  any differentiation of a particle method.
* Rollouts keep a checkpoint of the state every `every` steps and recompute
  each segment with autograd on the way back (reverse-in-time recomputation).

The step length is part of the differentiated function (the minimum of the
reference's dt rules, sph.h:545-568 and sphIncompress.cpp:118-140), as is the
time that decides the freeze hook; reinitialisation of density
(REINIT_DENS_EVERY_N_STEPS) is not supported and is rejected.
"""
from __future__ import annotations

import ctypes
import math
import time

import numpy as np
import torch

import sphdem as sd

STATE = ("r", "v", "vhat", "f", "fdrag", "dens", "h", "dddt", "dt", "time")
PARAMS = ("viscosity", "drag_a", "drag_b", "dem_k", "dem_gamma")
GRAVITY = -9.81


# ---------------------------------------------------------------- C ABI mirrors
class AdjConsts(ctypes.Structure):
    _fields_ = [(k, ctypes.c_double) for k in (
        "wcon", "viscosity", "dens", "pi", "dem_radius", "dem_vol", "dem_k", "dem_gamma",
        "drag_a", "drag_b", "coupling_radius", "porosity_incr")] + [
            ("shift_len", ctypes.c_double * 3)]


class AdjFields(ctypes.Structure):
    _fields_ = [(k, ctypes.c_void_p) for k in (
        "r", "vhat", "h", "dens", "mass", "omega", "pdr2", "fpfb", "fv", "por", "fdrag", "shep",
        "iam")]


class AdjGrads(ctypes.Structure):
    _fields_ = [(k, ctypes.c_void_p) for k in (
        "r", "vhat", "h", "dens", "omega", "pdr2", "fpfb", "fv", "por", "fdrag", "params")]


def _bind(lib):
    if getattr(lib, "_adj_bound", False):
        return lib
    vp, i64, i32 = ctypes.c_void_p, ctypes.c_int64, ctypes.c_int32
    pc, pf, pg = ctypes.POINTER(AdjConsts), ctypes.POINTER(AdjFields), ctypes.POINTER(AdjGrads)
    lib.sphdem_adj_force.argtypes = [i32, vp, vp, vp, vp, pf, pc, vp, vp, vp]
    lib.sphdem_adj_force_bwd.argtypes = [i32, vp, vp, vp, vp, pf, pc, vp, vp, pg, vp]
    for nm in ("sphdem_adj_omega", "sphdem_adj_dddt"):
        getattr(lib, nm).argtypes = [i32, vp, vp, vp, vp, pf, pc, vp, vp]
        getattr(lib, nm + "_bwd").argtypes = [i32, vp, vp, vp, vp, pf, pc, vp, pg, vp]
    lib.sphdem_adj_coupling.argtypes = [i32, vp, vp, vp, vp, pf, pc, vp, vp, vp, vp]
    lib.sphdem_adj_coupling_bwd.argtypes = [i32, vp, vp, vp, vp, pf, pc, vp, vp, vp, pg, vp]
    lib.sphdem_adj_coupling_bwd_noalloc.argtypes = [i32, vp, vp, vp, vp, pf, pc, vp, vp, i64,
                                                    pg, vp]
    lib.sphdem_adj_contacts.argtypes = [i64, vp, vp, vp, pf, pc, vp, vp]
    lib.sphdem_adj_contacts_bwd.argtypes = [i64, vp, vp, vp, pf, pc, vp, pg, vp]
    lib.sphdem_adj_fpi.argtypes = [i32, vp, vp, vp, vp, pf, pc, vp, vp]
    lib.sphdem_adj_fpi_bwd.argtypes = [i32, vp, vp, vp, vp, pf, pc, vp, pg, vp]
    lib._adj_bound = True
    return lib


def _adj_lib():
    """libsphdem.so with the adjoint entry points bound; raises if it is not
    built (there is no CPU fallback)."""
    return _bind(sd.load_library())


def _ptr(t):
    return None if t is None else ctypes.c_void_p(t.data_ptr())


def _stream():
    return ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)


# Per-kernel launch timing: when KERNEL_TIMES is a dict, each HIP pass launch is
# bracketed by CUDA events on the stream it runs on and appended to
# KERNEL_TIMES[name] as (start, end, work items); kernel_ms() reduces them.
KERNEL_TIMES = None


def _launch(name, items, fn):
    if KERNEL_TIMES is None:
        return fn()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record()
    rc = fn()
    e1.record()
    KERNEL_TIMES.setdefault(name, []).append((e0, e1, items))
    return rc


def kernel_ms() -> dict:
    """{name: (total ms, launches, total work items)} of the recorded launches."""
    torch.cuda.synchronize()
    out = {}
    for k, v in (KERNEL_TIMES or {}).items():
        out[k] = (sum(a.elapsed_time(b) for a, b, _ in v), len(v), sum(n for _, _, n in v))
    return out


# ---------------------------------------------------------------- pair sets
class PairSet:
    """One neighbour search: pairs (a, b) with b's periodic image shift, plus
    the CSR form grouped by a (for the per-query kernels)."""

    def __init__(self, abi, n: int, shift_len: torch.Tensor, symmetric: bool = False):
        """abi: an (m, 3) int tensor of (a, b, image) rows, or the three columns.
        symmetric: the set is known to hold (b, a, 26 - image) for every
        (a, b, image) -- the library's SPH search, whose criterion uses
        max(h_a, h_b) -- so require_symmetric() need not check it."""
        if isinstance(abi, torch.Tensor):
            abi = abi.to(torch.int32)
            abi = (abi[:, 0], abi[:, 1], abi[:, 2])
        self.n = n
        self.a, self.b, self.img = (t.to(torch.int32).contiguous() for t in abi)
        self.count = int(self.a.numel())
        if self.count:
            lo = int(torch.minimum(self.a.min(), self.b.min()))
            hi = int(torch.maximum(self.a.max(), self.b.max()))
            if lo < 0 or hi >= n or int(self.img.min()) < 0 or int(self.img.max()) > 26:
                raise ValueError("pair indices out of range")
        im = self.img.long()
        unit = torch.stack([im % 3 - 1, (im // 3) % 3 - 1, im // 9 - 1], dim=1)
        self.shift = unit.to(shift_len.dtype) * shift_len
        self._csr = None
        self._symmetric = symmetric or self.count == 0

    def require_symmetric(self):
        """Raises ValueError unless every (a, b, image) has its (b, a, 26 - image):
        the backward kernels of force, omega and dddt gather the partner terms
        from the query's own list and are correct only then.  Checked once per
        set (two sorts of the pair keys)."""
        if self._symmetric:
            return
        m = 27 * self.n
        a, b, im = self.a.long(), self.b.long(), self.img.long()
        fwd = torch.sort(a * m + b * 27 + im).values
        rev = torch.sort(b * m + a * 27 + (26 - im)).values
        if not torch.equal(fwd, rev):
            raise ValueError("the pair set is not symmetric: some (a, b, image) has no "
                             "(b, a, 26 - image); the gathered backward of force, omega and "
                             "dddt needs both")
        self._symmetric = True

    def csr(self, queries: torch.Tensor | None = None):
        """(query, off, b, img): one segment of pairs per query particle.

        The library's searches list each query's pairs contiguously, so the
        segments are read off as they come; a set where a query appears in
        two places is first sorted by a (stable).  queries: particles that
        must each get a segment (those without pairs get empty ones)."""
        key = None if queries is None else "all"
        if self._csr is None:
            self._csr = {}
        if key not in self._csr:
            a, b, img = self.a, self.b, self.img
            q, cnt = torch.unique_consecutive(a, return_counts=True)
            if q.numel() != torch.unique(q).numel():
                order = torch.sort(a.long(), stable=True).indices
                a, b, img = a[order], b[order], img[order]
                q, cnt = torch.unique_consecutive(a, return_counts=True)
            if queries is not None:
                seen = torch.zeros(self.n, dtype=torch.bool, device=a.device)
                seen[q.long()] = True
                allq = torch.zeros(self.n, dtype=torch.bool, device=a.device)
                allq[queries.long()] = True
                if bool((seen & ~allq).any()):
                    raise ValueError("pairs of a particle outside the query set")
                extra = torch.nonzero(allq & ~seen).flatten().to(q.dtype)
                q = torch.cat([q, extra])
                cnt = torch.cat([cnt, torch.zeros_like(extra, dtype=cnt.dtype)])
            off = torch.zeros(q.numel() + 1, dtype=torch.int64, device=a.device)
            off[1:] = torch.cumsum(cnt, 0)
            self._csr[key] = (q.to(torch.int32).contiguous(), off.contiguous(),
                              b.contiguous(), img.contiguous())
        return self._csr[key]


class LibraryPairs:
    """Pair sets of the library's own searches: a helper context holding the
    case is given positions and h, reset (CdataLL::reset) and asked for the
    pairs of each search kind (sphdem_get_pairs)."""

    # the library's SPH search is symmetric: its criterion uses max(h_a, h_b)
    # and one set of positions (its pair sets equal the oracle's brute force)
    symmetric_sph = True

    def __init__(self, case: sd.Case, device: int = 0):
        self.sim = sd.Sim(case, device=device)
        self.n = case.n
        self.device = device
        self.seconds = 0.0  # wall time spent in the searches
        self.cap = {}       # pair buffer size per search kind, grown on demand

    def __call__(self, r: torch.Tensor, h: torch.Tensor, kinds) -> dict:
        t0 = time.perf_counter()
        out = self._search(r, h, kinds)
        self.seconds += time.perf_counter() - t0
        return out

    def _search(self, r, h, kinds):
        """r (n, 3) and h (n) in original order, on the helper's device: set
        device to device (sphdem_set_field_device), reset, export."""
        rr = r.detach().to(torch.float64).contiguous()
        hh = h.detach().to(torch.float64).contiguous()
        if (not rr.is_cuda or rr.device.index != self.device or tuple(rr.shape) != (self.n, 3)
                or tuple(hh.shape) != (self.n,) or hh.device != rr.device):
            raise ValueError(f"positions {tuple(rr.shape)} and h {tuple(hh.shape)} must be "
                             f"({self.n}, 3) and ({self.n},) on cuda:{self.device}")
        torch.cuda.synchronize(rr.device)  # rr, hh complete before the library's stream reads
        self.sim.set_device("r", rr.data_ptr())
        self.sim.set_device("h", hh.data_ptr())
        self.sim.reset()
        if self.sim.num_deleted():
            raise sd.SphDemError("a particle left the domain during the rollout")
        return {k: self._pairs(k, rr.device) for k in kinds}

    def _pairs(self, kind, device):
        """One search's pairs written straight into device tensors
        (sphdem_get_pairs_device): one call that counts and writes into
        buffers of the kind's last size; a second one, with buffers of the
        count it returned, when those were too small."""
        lib, cnt = self.sim.lib, ctypes.c_int64()
        cap = self.cap.get(kind, 0)
        for _ in range(2):
            cols = [torch.empty(max(cap, 1), dtype=torch.int32, device=device) for _ in range(3)]
            torch.cuda.synchronize(device)  # the buffers exist before the library's stream writes
            rc = lib.sphdem_get_pairs_device(self.sim.ctx, kind, ctypes.byref(cnt),
                                             *[ctypes.c_void_p(c.data_ptr()) for c in cols], cap)
            m = cnt.value
            if rc == sd.E_CAPACITY and m > cap:
                cap = m + m // 16 + 1024
                self.cap[kind] = cap
                continue
            self.sim._check(rc)
            return tuple(c[:m] for c in cols)
        raise sd.SphDemError(f"pair export of kind {kind}: {m} pairs after a buffer of {cap}")

    def close(self):
        self.sim.close()


# ---------------------------------------------------------------- Wendland
def kK(c, q):
    t = 2.0 - q
    return torch.where(q <= 2.0, c * t ** 4 * (1.0 + 2.0 * q), torch.zeros_like(q))


def kW(c, q, h):
    return kK(c, q) / h ** 3


def kF(c, q, h):
    t = 2.0 - q
    return torch.where(q <= 2.0, -10.0 * c * t ** 3 / h ** 5, torch.zeros_like(q))


def kPdWh(c, q, h):
    t = 2.0 - q
    K = c * t ** 4 * (1.0 + 2.0 * q)
    return torch.where(q <= 2.0, (10.0 * c * q * q * t ** 3 - 3.0 * K) / h ** 4,
                       torch.zeros_like(q))


def _g(t, idx):
    """t[idx] along dim 0 as index_select, whose backward is a scatter-add
    (advanced indexing's backward sorts the indices)."""
    return t.index_select(0, idx)


def _dx(ps: PairSet, r):
    a, b = ps.a.long(), ps.b.long()
    dx = _g(r, a) - (_g(r, b) + ps.shift)
    return a, b, dx, torch.sqrt((dx * dx).sum(1))


# ---------------------------------------------------------------- HIP passes
class _Ctx:
    """What the pair passes need besides the differentiable tensors."""

    def __init__(self, model, th):
        self.m = model
        self.consts = model.consts(th)


def _fields(C, **kw):
    f = AdjFields()
    f.mass = C.m.mass.data_ptr()
    f.iam = C.m.iam32.data_ptr()
    for k, v in kw.items():
        setattr(f, k, None if v is None else v.data_ptr())
    return f


def _zeros_like_all(*ts):
    return [torch.zeros_like(t) for t in ts]


def _lib_ok(rc, what):
    if rc != 0:
        raise sd.SphDemError(f"{what} failed ({rc})")


class _ForceFn(torch.autograd.Function):
    @staticmethod
    def forward(ctx, r, vhat, h, dens, omega, pdr2, visc, ps, C):
        lib = _adj_lib()
        fv, fp = torch.zeros_like(r), torch.zeros_like(r)
        cs = C.m.consts_with(C.consts, viscosity=float(visc.detach()))
        fl = _fields(C, r=r, vhat=vhat, h=h, dens=dens, omega=omega, pdr2=pdr2)
        q, off, b, img = ps.csr()
        _lib_ok(_launch("force", ps.count, lambda: lib.sphdem_adj_force(
            int(q.numel()), _ptr(q), _ptr(off), _ptr(b), _ptr(img), ctypes.byref(fl),
            ctypes.byref(cs), _ptr(fv), _ptr(fp), _stream())), "sphdem_adj_force")
        ctx.save_for_backward(r, vhat, h, dens, omega, pdr2, visc)
        ctx.ps, ctx.C, ctx.cs = ps, C, cs
        return fv, fp

    @staticmethod
    def backward(ctx, gfv, gfp):
        r, vhat, h, dens, omega, pdr2, visc = ctx.saved_tensors
        gfv = torch.zeros_like(r) if gfv is None else gfv.contiguous()
        gfp = torch.zeros_like(r) if gfp is None else gfp.contiguous()
        gr, gv, gh, gd, go, gp = _zeros_like_all(r, vhat, h, dens, omega, pdr2)
        par = torch.zeros(5, dtype=r.dtype, device=r.device)
        fl = _fields(ctx.C, r=r, vhat=vhat, h=h, dens=dens, omega=omega, pdr2=pdr2)
        g = AdjGrads(r=gr.data_ptr(), vhat=gv.data_ptr(), h=gh.data_ptr(), dens=gd.data_ptr(),
                     omega=go.data_ptr(), pdr2=gp.data_ptr(), params=par.data_ptr())
        ps = ctx.ps
        q, off, b, img = ps.csr()
        _lib_ok(_launch("force_bwd", ps.count, lambda: _adj_lib().sphdem_adj_force_bwd(
            int(q.numel()), _ptr(q), _ptr(off), _ptr(b), _ptr(img), ctypes.byref(fl),
            ctypes.byref(ctx.cs), _ptr(gfv), _ptr(gfp), ctypes.byref(g), _stream())),
            "sphdem_adj_force_bwd")
        return gr, gv, gh, gd, go, gp, par[0].reshape(visc.shape), None, None


def _csr_call(name, ps, *args, queries=None):
    q, off, b, img = ps.csr(queries)
    fn = getattr(_adj_lib(), "sphdem_adj_" + name)
    _lib_ok(_launch(name, ps.count, lambda: fn(int(q.numel()), _ptr(q), _ptr(off), _ptr(b),
                                                 _ptr(img), *args, _stream())),
            "sphdem_adj_" + name)


class _OmegaFn(torch.autograd.Function):
    @staticmethod
    def forward(ctx, r, h, dens, ps, C):
        out = torch.ones_like(h)
        fl = _fields(C, r=r, h=h, dens=dens)
        # every SPH-group particle, neighbours or not (calcOmega's self term)
        _csr_call("omega", ps, ctypes.byref(fl), ctypes.byref(C.consts), _ptr(out),
                  queries=C.m.grp_idx)
        ctx.save_for_backward(r, h, dens)
        ctx.ps, ctx.C = ps, C
        return out

    @staticmethod
    def backward(ctx, go):
        r, h, dens = ctx.saved_tensors
        gr, gh, gd = _zeros_like_all(r, h, dens)
        fl = _fields(ctx.C, r=r, h=h, dens=dens)
        g = AdjGrads(r=gr.data_ptr(), h=gh.data_ptr(), dens=gd.data_ptr())
        _csr_call("omega_bwd", ctx.ps, ctypes.byref(fl), ctypes.byref(ctx.C.consts),
                  _ptr(go.contiguous()), ctypes.byref(g), queries=ctx.C.m.grp_idx)
        return gr, gh, gd, None, None


class _DddtFn(torch.autograd.Function):
    @staticmethod
    def forward(ctx, r, vhat, h, omega, ps, C):
        out = torch.zeros_like(h)
        fl = _fields(C, r=r, vhat=vhat, h=h, omega=omega)
        _csr_call("dddt", ps, ctypes.byref(fl), ctypes.byref(C.consts), _ptr(out))
        ctx.save_for_backward(r, vhat, h, omega)
        ctx.ps, ctx.C = ps, C
        return out

    @staticmethod
    def backward(ctx, gd_):
        r, vhat, h, omega = ctx.saved_tensors
        gr, gv, gh, go = _zeros_like_all(r, vhat, h, omega)
        fl = _fields(ctx.C, r=r, vhat=vhat, h=h, omega=omega)
        g = AdjGrads(r=gr.data_ptr(), vhat=gv.data_ptr(), h=gh.data_ptr(), omega=go.data_ptr())
        _csr_call("dddt_bwd", ctx.ps, ctypes.byref(fl), ctypes.byref(ctx.C.consts),
                  _ptr(gd_.contiguous()), ctypes.byref(g))
        return gr, gv, gh, go, None, None


class _CouplingFn(torch.autograd.Function):
    @staticmethod
    def forward(ctx, r, dens, fdrag, ps, C):
        por = torch.zeros_like(dens)
        shep = torch.zeros_like(dens)
        drag = torch.zeros_like(r)
        fl = _fields(C, r=r, dens=dens, fdrag=fdrag)
        _csr_call("coupling", ps, ctypes.byref(fl), ctypes.byref(C.consts), _ptr(por),
                  _ptr(shep), _ptr(drag))
        ctx.save_for_backward(r, dens, fdrag, shep)
        ctx.ps, ctx.C = ps, C
        return por, drag

    @staticmethod
    def backward(ctx, gpor, gdrag):
        r, dens, fdrag, shep = ctx.saved_tensors
        gpor = torch.zeros_like(dens) if gpor is None else gpor.contiguous()
        gdrag = torch.zeros_like(r) if gdrag is None else gdrag.contiguous()
        gr, gd, gf = _zeros_like_all(r, dens, fdrag)
        work = torch.zeros_like(dens)
        fl = _fields(ctx.C, r=r, dens=dens, fdrag=fdrag, shep=shep)
        g = AdjGrads(r=gr.data_ptr(), dens=gd.data_ptr(), fdrag=gf.data_ptr())
        _csr_call("coupling_bwd", ctx.ps, ctypes.byref(fl), ctypes.byref(ctx.C.consts),
                  _ptr(gpor), _ptr(gdrag), _ptr(work), ctypes.byref(g))
        return gr, gd, gf, None, None


class _FpiFn(torch.autograd.Function):
    @staticmethod
    def forward(ctx, r, vhat, h, dens, fpfb, fv, por, visc, drag_a, drag_b, ps, C):
        lib = _adj_lib()
        out = torch.zeros_like(r)
        cs = C.m.consts_with(C.consts, viscosity=float(visc.detach()),
                             drag_a=float(drag_a.detach()), drag_b=float(drag_b.detach()))
        q, off, b, img = ps.csr()
        fl = _fields(C, r=r, vhat=vhat, h=h, dens=dens, fpfb=fpfb, fv=fv, por=por)
        _lib_ok(_launch("fpi", ps.count, lambda: lib.sphdem_adj_fpi(
            int(q.numel()), _ptr(q), _ptr(off), _ptr(b), _ptr(img), ctypes.byref(fl),
            ctypes.byref(cs), _ptr(out), _stream())), "sphdem_adj_fpi")
        ctx.save_for_backward(r, vhat, h, dens, fpfb, fv, por, visc, drag_a, drag_b)
        ctx.ps, ctx.C, ctx.cs = ps, C, cs
        return out

    @staticmethod
    def backward(ctx, gout):
        r, vhat, h, dens, fpfb, fv, por, visc, drag_a, drag_b = ctx.saved_tensors
        gout = gout.contiguous()
        gr, gv, gh, gd, gpf, gfv, gpo = _zeros_like_all(r, vhat, h, dens, fpfb, fv, por)
        par = torch.zeros(5, dtype=r.dtype, device=r.device)
        fl = _fields(ctx.C, r=r, vhat=vhat, h=h, dens=dens, fpfb=fpfb, fv=fv, por=por)
        g = AdjGrads(r=gr.data_ptr(), vhat=gv.data_ptr(), h=gh.data_ptr(), dens=gd.data_ptr(),
                     fpfb=gpf.data_ptr(), fv=gfv.data_ptr(), por=gpo.data_ptr(),
                     params=par.data_ptr())
        q, off, b, img = ctx.ps.csr()
        _lib_ok(_launch("fpi_bwd", ctx.ps.count, lambda: _adj_lib().sphdem_adj_fpi_bwd(
            int(q.numel()), _ptr(q), _ptr(off), _ptr(b), _ptr(img), ctypes.byref(fl),
            ctypes.byref(ctx.cs), _ptr(gout), ctypes.byref(g), _stream())), "sphdem_adj_fpi_bwd")
        return (gr, gv, gh, gd, gpf, gfv, gpo, par[0].reshape(visc.shape),
                par[1].reshape(drag_a.shape), par[2].reshape(drag_b.shape), None, None)


class _ContactsFn(torch.autograd.Function):
    @staticmethod
    def forward(ctx, r, vhat, dem_k, dem_gamma, ps, C):
        lib = _adj_lib()
        out = torch.zeros_like(r)
        cs = C.m.consts_with(C.consts, dem_k=float(dem_k.detach()),
                             dem_gamma=float(dem_gamma.detach()))
        fl = _fields(C, r=r, vhat=vhat)
        _lib_ok(_launch("contacts", ps.count, lambda: lib.sphdem_adj_contacts(
            ps.count, _ptr(ps.a), _ptr(ps.b), _ptr(ps.img), ctypes.byref(fl), ctypes.byref(cs),
            _ptr(out), _stream())), "sphdem_adj_contacts")
        ctx.save_for_backward(r, vhat, dem_k, dem_gamma)
        ctx.ps, ctx.C, ctx.cs = ps, C, cs
        return out

    @staticmethod
    def backward(ctx, gout):
        r, vhat, dem_k, dem_gamma = ctx.saved_tensors
        gout = gout.contiguous()
        gr, gv = _zeros_like_all(r, vhat)
        par = torch.zeros(5, dtype=r.dtype, device=r.device)
        fl = _fields(ctx.C, r=r, vhat=vhat)
        g = AdjGrads(r=gr.data_ptr(), vhat=gv.data_ptr(), params=par.data_ptr())
        ps = ctx.ps
        _lib_ok(_launch("contacts_bwd", ps.count, lambda: _adj_lib().sphdem_adj_contacts_bwd(
            ps.count, _ptr(ps.a), _ptr(ps.b), _ptr(ps.img), ctypes.byref(fl),
            ctypes.byref(ctx.cs), _ptr(gout), ctypes.byref(g), _stream())),
            "sphdem_adj_contacts_bwd")
        return gr, gv, par[3].reshape(dem_k.shape), par[4].reshape(dem_gamma.shape), None, None


def _dev(*ts):
    for t in ts:
        if not (t.is_cuda and t.dtype == torch.float64 and t.is_contiguous()):
            raise ValueError("the HIP passes take contiguous float64 tensors on the GPU")


class HipPasses:
    """Every pair pass of the step as a HIP kernel with a hand-written
    backward kernel."""

    def __init__(self):
        _adj_lib()

    # force, omega and dddt: their backward kernels gather over the query's
    # own list, so the set must be symmetric (checked before any launch)
    def omega(self, C, ps, r, h, dens):
        ps.require_symmetric()
        _dev(r, h, dens)
        return _OmegaFn.apply(r, h, dens, ps, C)

    def dddt(self, C, ps, r, vhat, h, omega):
        ps.require_symmetric()
        _dev(r, vhat, h, omega)
        return _DddtFn.apply(r, vhat, h, omega, ps, C)

    def coupling(self, C, ps, r, dens, fdrag):
        _dev(r, dens, fdrag)
        return _CouplingFn.apply(r, dens, fdrag, ps, C)

    def force(self, C, ps, r, vhat, h, dens, omega, pdr2, visc):
        ps.require_symmetric()
        _dev(r, vhat, h, dens, omega, pdr2)
        return _ForceFn.apply(r, vhat, h, dens, omega, pdr2, visc, ps, C)

    def fpi(self, C, ps, r, vhat, h, dens, fpfb, fv, por, visc, drag_a, drag_b):
        _dev(r, vhat, h, dens, fpfb, fv, por)
        return _FpiFn.apply(r, vhat, h, dens, fpfb, fv, por, visc, drag_a, drag_b, ps, C)

    def contacts(self, C, ps, r, vhat, dem_k, dem_gamma):
        _dev(r, vhat)
        return _ContactsFn.apply(r, vhat, dem_k, dem_gamma, ps, C)


# ---------------------------------------------------------------- the step
def _morton_order(r: np.ndarray, P, edge: float) -> np.ndarray:
    """Stable order of the particles by the Morton code of their cells (edge
    `edge`, from the domain's lower corner)."""
    lo = np.array([P.proc_domain[0], P.proc_domain[2], P.proc_domain[4]])
    c = np.clip(np.floor((r - lo) / edge), 0, 1023).astype(np.uint64)

    def spread(x):
        x = (x | (x << np.uint64(16))) & np.uint64(0x030000FF)
        x = (x | (x << np.uint64(8))) & np.uint64(0x0300F00F)
        x = (x | (x << np.uint64(4))) & np.uint64(0x030C30C3)
        x = (x | (x << np.uint64(2))) & np.uint64(0x09249249)
        return x

    key = spread(c[:, 0]) | (spread(c[:, 1]) << np.uint64(1)) | (spread(c[:, 2]) << np.uint64(2))
    return np.argsort(key, kind="stable")


SEDIMENT_FLAGS = sd.F_VISC_MORRIS | sd.F_VAR_H_CORRECTION | sd.F_LIQ_DEM | sd.F_CUSTOM_WALL


class StepModel:
    """The FP64 step of the sediment3D case as a differentiable function.

    passes: HipPasses (the default) or an object with the same three methods.
    pairs:  callable (r, h, kinds) -> {kind: (m, 3) int tensor of (a, b, image)}.
    """

    def __init__(self, case: sd.Case, device="cuda", passes=None, pairs=None):
        P = case.params
        if P.flags != SEDIMENT_FLAGS:
            raise ValueError("the adjoint restates the sediment3D flag set only "
                             "(VISC_MORRIS, VAR_H_CORRECTION, LIQ_DEM, "
                             "LIQ_DEM_CUSTOM_WALL_CONTACTS)")
        if case.hookset != sd.HOOKS_SEDIMENT:
            raise ValueError("the adjoint restates the sediment3D hooks only")
        if P.reinit_every < 2 ** 31 - 1 or P.reinit_at >= 0:
            raise ValueError("density reinitialisation is not supported by the adjoint")
        if P.damptime > 0:
            raise ValueError("velocity damping (DAMPTIME) is not supported by the adjoint")
        iam = np.asarray(case.iam)
        if not np.isin(iam, (sd.SPH, sd.SPH_BOUNDARY, sd.DEM)).all():
            raise ValueError("particle types other than sph, sphBoundary and dem")
        self.case = case
        self.P = P
        self.device = torch.device(device)
        dev, f64 = self.device, torch.float64
        self.n = case.n
        # The model keeps its particles in Morton order of the cells of the
        # initial positions (cell edge 2H), so that the pair kernels' partner
        # reads stay local; perm[i] is the original index of row i, inv its
        # inverse.  Inputs, pair sets, the loss and the gradients are mapped at
        # the boundary; states returned by step() are in this order
        # (original() maps a per-particle tensor back).
        perm = _morton_order(np.asarray(case.r), P, 2.0 * P.h)
        self.perm = torch.as_tensor(perm, dtype=torch.long, device=dev)
        self.inv = torch.empty_like(self.perm)
        self.inv[self.perm] = torch.arange(self.n, device=dev)
        iam = iam[perm]
        it = torch.as_tensor(iam, device=dev)
        self.iam32 = it.to(torch.int32).contiguous()
        self.sph = it == sd.SPH
        self.sphb = it == sd.SPH_BOUNDARY
        self.grp = self.sph | self.sphb
        self.dem = it == sd.DEM
        self.grp_idx = torch.nonzero(self.grp).flatten().to(torch.int32)
        self.mass = torch.as_tensor(np.asarray(case.mass)[perm], dtype=f64, device=dev).contiguous()
        dem_orig = np.nonzero(np.asarray(case.iam) == sd.DEM)[0]
        self.dem_rows = self.inv[torch.as_tensor(dem_orig, dtype=torch.long, device=dev)]
        self.shift_len = torch.tensor([(P.rmax[i] - P.rmin[i]) if P.periodic[i] else 0.0
                                       for i in range(3)], dtype=f64, device=dev)
        self.passes = HipPasses() if passes is None else passes
        self.pairs_fn = pairs
        self.cache = {}

    # ---- constants
    def consts(self, th) -> AdjConsts:
        P = self.P
        c = AdjConsts(wcon=P.wcon, viscosity=P.viscosity, dens=P.dens, pi=P.pi,
                      dem_radius=P.dem_radius, dem_vol=P.dem_vol, dem_k=P.dem_k,
                      dem_gamma=P.dem_gamma, drag_a=P.drag_a, drag_b=P.drag_b,
                      coupling_radius=P.coupling_radius, porosity_incr=P.porosity_incr)
        for i in range(3):
            c.shift_len[i] = float(self.shift_len[i])
        return c

    @staticmethod
    def consts_with(c: AdjConsts, **kw) -> AdjConsts:
        out = AdjConsts()
        ctypes.pointer(out)[0] = c
        for k, v in kw.items():
            setattr(out, k, v)
        return out

    def params(self, requires_grad=True) -> dict:
        """The differentiable constants as 0-d float64 tensors."""
        return {k: torch.tensor(getattr(self.P, k), dtype=torch.float64, device=self.device,
                                requires_grad=requires_grad) for k in PARAMS}

    def initial_state(self, v_init=None) -> dict:
        """The state after upload: r, v = vhat (v_init, default the case's),
        zero f, fdrag and dddt, dt = 0, time = 0."""
        c, dev, f64 = self.case, self.device, torch.float64
        t = lambda a: self.from_original(torch.as_tensor(np.ascontiguousarray(a), dtype=f64,
                                                         device=dev))
        v = t(c.v) if v_init is None else self.from_original(v_init)
        vh = t(c.vhat) if v_init is None else v
        z3 = torch.zeros((self.n, 3), dtype=f64, device=dev)
        return {"r": t(c.r), "v": v, "vhat": vh, "f": z3, "fdrag": z3.clone(), "dens": t(c.dens),
                "h": t(c.h), "dddt": torch.zeros(self.n, dtype=f64, device=dev),
                "dt": torch.zeros((), dtype=f64, device=dev),
                "time": torch.zeros((), dtype=f64, device=dev)}

    def from_original(self, t: torch.Tensor) -> torch.Tensor:
        """A per-particle tensor in original index order -> the model's order."""
        return t.index_select(0, self.perm)

    def original(self, t: torch.Tensor) -> torch.Tensor:
        """A per-particle tensor in the model's order -> original index order."""
        return t.index_select(0, self.inv)

    # ---- pieces
    def _pairs(self, key, r, h, kinds):
        """The pair sets of a reset at positions r, h (model order), searched in
        original order and renumbered to the model's rows."""
        if key not in self.cache:
            got = self.pairs_fn(self.original(r), self.original(h), kinds)
            out = {}
            for k in kinds:
                g = got[k]
                cols = (g[:, 0], g[:, 1], g[:, 2]) if isinstance(g, torch.Tensor) else g
                a = self.inv.index_select(0, cols[0].long())
                b = self.inv.index_select(0, cols[1].long())
                sym = k == sd.PAIRS_SPH and getattr(self.pairs_fn, "symmetric_sph", False)
                out[k] = PairSet((a, b, cols[2]), self.n, self.shift_len, symmetric=sym)
            self.cache[key] = out
        return self.cache[key]

    def _wrap(self, r):
        """Periodic wrap of CdataLL::updateDomain (dataLL.cpp:370-403, 469-484)."""
        P = self.P
        cols = []
        for i in range(3):
            x = r[:, i]
            if P.periodic[i]:
                dmin = P.proc_domain[2 * i]
                dsp = P.proc_domain[2 * i + 1] - dmin
                oc = (x - dmin) / dsp + 1.0
                x = torch.where(oc < 1.0, (P.rmax[i] - P.rmin[i]) + x,
                                torch.where(oc >= 2.0, (P.rmin[i] - P.rmax[i]) + x, x))
            cols.append(x)
        return torch.stack(cols, dim=1)

    def _floor(self, r, vhat, f, th):
        """addCustomBoundaries (sediment3D/customSim.cpp:67-81): the z = RMIN
        floor as a LINEAR spring-dashpot on DEM particles."""
        P = self.P
        ov = P.rmin[2] + P.dem_radius - r[:, 2]
        on = self.dem & (ov > 0)
        s = th["dem_k"] * ov + th["dem_gamma"] * (-vhat[:, 2])
        fz = f[:, 2] + torch.where(on, s / self.mass, torch.zeros_like(s))
        return torch.stack([f[:, 0], f[:, 1], fz], dim=1)

    def _new_dt(self, ps, h, f, th, time):
        """The dt minimum of middle(): maxdt, the viscous-pair rules of
        calcViscForce (sph.h:545-560) over pairs with an SPH first member,
        LIQ_DEM_TIMESTEP, DEM_TIMESTEP after the drop time, and accelCondition
        (sph.h:566-568) over the SPH group and DEM."""
        P = self.P
        terms = [torch.tensor(P.maxdt, dtype=h.dtype, device=h.device)]
        a, b = ps.a.long(), ps.b.long()
        sel = _g(self.sph, a)
        if bool(sel.any()):
            # the smallest hav over the pairs, found without a graph over every
            # pair, then recomputed with its gradient from its two particles
            with torch.no_grad():
                hs = 0.5 * (_g(h, a) + _g(h, b))
                k = torch.where(sel, hs, torch.full_like(hs, math.inf)).argmin()
            mh = 0.5 * (h[a[k]] + h[b[k]])
            terms.append(0.6 * mh / (2.0 * P.spsound))
            terms.append(0.125 * mh ** 2 / th["viscosity"])
        terms.append(torch.tensor(P.liq_dem_timestep, dtype=h.dtype, device=h.device))
        if float(time.detach()) > P.time_drop:
            terms.append(torch.tensor(P.dem_timestep, dtype=h.dtype, device=h.device))
        f2 = (f * f).sum(1)
        on = (self.grp | self.dem) & (f2 > 0)
        if bool(on.any()):  # |f| = 0 sets no bound (and would give 0/0 adjoints)
            terms.append((0.25 * torch.sqrt(h[on] / torch.sqrt(f2[on]))).min())
        return torch.stack(terms).min()

    def _freeze(self, S, time):
        """freezeParticles (sediment3D/customSim.cpp:12-25) on DEM while
        time <= TIME_DROP_PARTICLE."""
        if float(time.detach()) <= self.P.time_drop:
            keep = (~self.dem)[:, None]
            for k in ("f", "v", "vhat"):
                S[k] = torch.where(keep, S[k], torch.zeros_like(S[k]))

    # ---- one step (run.cpp:102-118 with the sediment3D hooks)
    def step(self, S: dict, th: dict, k: int) -> dict:
        P, pz = self.P, self.passes
        grp, sph, dem = self.grp[:, None], self.sph[:, None], self.dem[:, None]
        m = self.mass
        C = _Ctx(self, th)
        r, v, vhat, f = S["r"], S["v"], S["vhat"], S["f"]
        dens, h, dddt, fdrag = S["dens"], S["h"], S["dddt"], S["fdrag"]
        dt, time = S["dt"], S["time"]
        # start (sphIncompress.cpp:3-38)
        half = dt / 2
        r = torch.where(grp | dem, r + half * vhat, r)
        v0, vhat0 = v, vhat
        v = torch.where(sph, v + half * f, v)
        vhat = torch.where(sph, vhat + half * f, vhat)
        dens = torch.where(self.grp, dens + half * dddt, dens)
        h = torch.where(self.grp, P.hfac * (m / dens) ** (1.0 / 3.0), h)
        time = time + half
        f = torch.where(grp, torch.zeros_like(f), f)
        gz = torch.zeros_like(f)
        gz[:, 2] = GRAVITY
        f = torch.where(sph, f + gz, f)                       # gravity on sph
        # middle (sphIncompress.cpp:40-165)
        r = self._wrap(r)
        pm = self._pairs((k, "middle"), r, h, (sd.PAIRS_SPH, sd.PAIRS_SPH2DEM))
        por_sum, drag_s = pz.coupling(C, pm[sd.PAIRS_SPH2DEM], r, dens, fdrag)
        por = torch.where(self.grp, torch.clamp(1.0 - por_sum, min=0.37),
                          torch.ones_like(por_sum))         # finalisePorosity sph.h:1289-1296
        f = torch.where(grp, f + drag_s, f)
        omega = pz.omega(C, pm[sd.PAIRS_SPH], r, h, dens)
        press = P.prb * ((dens / (P.refd * por)) ** P.gamma - 1.0)
        pdr2 = torch.where(self.grp, press / (dens * dens), torch.zeros_like(press))
        fv, fp = pz.force(C, pm[sd.PAIRS_SPH], r, vhat, h, dens, omega, pdr2, th["viscosity"])
        f = f + fv + fp
        new_dt = self._new_dt(pm[sd.PAIRS_SPH], h, f, th, time)
        old_dt, dt = dt, new_dt
        kdt = 0.5 * old_dt + 0.5 * dt
        v = torch.where(sph, v0 + kdt * f, v)
        vhat = torch.where(sph, vhat0 + kdt * f, vhat)
        dddt = torch.where(self.grp, torch.zeros_like(dddt), dddt)
        half = dt / 2
        r = torch.where(dem, r + half * vhat, r)
        v0d, vhat0d = v, vhat
        v = torch.where(dem, v + half * f, v)
        vhat = torch.where(dem, vhat + half * f, vhat)
        f = torch.where(dem, torch.zeros_like(f), f)
        # beforeEnd hooks (sediment3D/customSim.cpp:97-107)
        f = torch.where(dem, f + gz, f)
        f = self._floor(r, vhat, f, th)
        T = {"f": f, "v": v, "vhat": vhat}
        self._freeze(T, time)
        f, v, vhat = T["f"], T["v"], T["vhat"]
        # end (sphIncompress.cpp:167-251)
        r = torch.where(grp, r + half * vhat, r)
        time = time + half
        r = self._wrap(r)
        pe = self._pairs((k, "end"), r, h, (sd.PAIRS_SPH, sd.PAIRS_DEM2SPH, sd.PAIRS_DEM))
        omega = pz.omega(C, pe[sd.PAIRS_SPH], r, h, dens)
        dddt = torch.where(self.grp, pz.dddt(C, pe[sd.PAIRS_SPH], r, vhat, h, omega), dddt)
        fd = pz.fpi(C, pe[sd.PAIRS_DEM2SPH], r, vhat, h, dens, fp, fv, por, th["viscosity"],
                    th["drag_a"], th["drag_b"])
        f = torch.where(dem, f + fd, f)
        fdrag = torch.where(dem, fd, drag_s)
        f = f + pz.contacts(C, pe[sd.PAIRS_DEM], r, vhat, th["dem_k"], th["dem_gamma"])
        v = torch.where(dem, v0d + dt * f, v)
        vhat = torch.where(dem, vhat0d + dt * f, vhat)
        dens = torch.where(self.grp, dens + half * dddt, dens)
        h = torch.where(self.grp, P.hfac * (m / dens) ** (1.0 / 3.0), h)
        T = {"f": f, "v": v, "vhat": vhat}
        self._freeze(T, time)                                  # afterEnd
        return {"r": r, "v": T["v"], "vhat": T["vhat"], "f": T["f"], "fdrag": fdrag,
                "dens": dens, "h": h, "dddt": dddt, "dt": dt, "time": time}

    # ---- rollouts
    def clear_pairs(self):
        self.cache.clear()

    def rollout(self, S, th, steps: int, first: int = 0) -> dict:
        """Steps first .. first+steps-1 from state S.  Pair sets are cached by
        step index (clear_pairs() before a rollout of other inputs)."""
        for k in range(first, first + steps):
            S = self.step(S, th, k)
        return S

    def dem_loss(self, S, target: torch.Tensor) -> torch.Tensor:
        """Mean squared DEM z-position against a target profile."""
        z = S["r"][:, 2].index_select(0, self.dem_rows)   # DEM in original order
        return ((z - target) ** 2).mean()

    def gradient(self, steps: int, target: torch.Tensor, v_init=None, every: int = 8):
        """Loss after `steps` steps and its gradient with respect to PARAMS and
        the initial velocity (v = vhat = v_init), by reverse mode with a state
        checkpoint every `every` steps and recomputation of each segment."""
        if v_init is None:
            v_init = torch.as_tensor(self.case.v, dtype=torch.float64, device=self.device)
        th = self.params(requires_grad=True)
        v0 = v_init.detach().clone().contiguous().requires_grad_(True)
        self.cache.clear()
        bounds = list(range(0, steps, every)) + [steps]
        ckpt = []
        with torch.no_grad():
            S = self.initial_state(v0.detach())
            for s0, s1 in zip(bounds[:-1], bounds[1:]):
                ckpt.append({k: x.clone() for k, x in S.items()})
                S = self.rollout(S, th, s1 - s0, s0)
            loss = float(self.dem_loss(S, target))
        gS = None
        for seg in reversed(range(len(ckpt))):
            s0, s1 = bounds[seg], bounds[seg + 1]
            if seg == 0:
                Sin = self.initial_state(v0)
                leaves = []
            else:
                Sin = {k: x.clone().requires_grad_(True) for k, x in ckpt[seg].items()}
                leaves = [Sin[k] for k in STATE]
            Sout = self.rollout(Sin, th, s1 - s0, s0)
            if gS is None:
                obj = self.dem_loss(Sout, target)
            else:
                obj = sum((Sout[k] * gS[k]).sum() for k in STATE if Sout[k].requires_grad)
            obj.backward()
            if seg > 0:
                gS = {k: (x.grad if x.grad is not None else torch.zeros_like(x))
                      for k, x in zip(STATE, leaves)}
            del Sout, obj
        grads = {k: float(th[k].grad) if th[k].grad is not None else 0.0 for k in PARAMS}
        grads["v_init"] = v0.grad if v0.grad is not None else torch.zeros_like(v0)
        return loss, grads
