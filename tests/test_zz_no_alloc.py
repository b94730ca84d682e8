"""sphdem_adj_coupling_bwd_noalloc (the coupling backward with its Shepard-sum
adjoints in the library's device-resident scratch array) against
sphdem_adj_coupling_bwd with a caller-given work buffer: two calls in a row,
the second on a smaller case, through the C ABI."""
import ctypes

import numpy as np
import pytest
import torch

import sphdem as sd
import sphdem_adjoint as sa


def normwise(a, b):
    a, b = np.asarray(a), np.asarray(b)
    return float(np.abs(a - b).max() / max(np.abs(b).max(), 1e-300))


def _coupling_inputs(case, seed):
    """A model state two steps in, its SPH->DEM pair set and seeded adjoints."""
    pairs = sa.LibraryPairs(case)
    model = sa.StepModel(case, device="cuda", pairs=pairs)
    th = model.params(requires_grad=False)
    S = model.initial_state()
    with torch.no_grad():
        S = model.rollout(S, th, 2)
    C = sa._Ctx(model, th)
    r, h = S["r"].contiguous(), S["h"].contiguous()
    ps = model._pairs(("noalloc", seed), r, h, (sd.PAIRS_SPH2DEM,))[sd.PAIRS_SPH2DEM]
    assert ps.count > 0
    gen = torch.Generator(device="cuda").manual_seed(seed)
    rnd = lambda t: torch.rand(t.shape, generator=gen, dtype=t.dtype, device=t.device) - 0.5
    dens = S["dens"].contiguous()
    fdrag = torch.where(model.dem[:, None], rnd(r), torch.zeros_like(r)).contiguous()
    lib = sa._adj_lib()
    q, off, b, img = ps.csr()
    por = torch.zeros_like(dens)
    shep = torch.zeros_like(dens)
    drag = torch.zeros_like(r)
    fl = sa._fields(C, r=r, dens=dens, fdrag=fdrag)
    rc = lib.sphdem_adj_coupling(int(q.numel()), sa._ptr(q), sa._ptr(off), sa._ptr(b),
                                 sa._ptr(img), ctypes.byref(fl), ctypes.byref(C.consts),
                                 sa._ptr(por), sa._ptr(shep), sa._ptr(drag), sa._stream())
    assert rc == 0
    gpor = rnd(dens).contiguous()
    gdrag = rnd(r).contiguous()
    return dict(C=C, pairs=pairs, q=q, off=off, b=b, img=img, r=r, dens=dens, fdrag=fdrag,
                shep=shep, gpor=gpor, gdrag=gdrag, n=case.n)


def _backward(x, noalloc):
    lib = sa._adj_lib()
    gr = torch.zeros_like(x["r"])
    gd = torch.zeros_like(x["dens"])
    gf = torch.zeros_like(x["fdrag"])
    g = sa.AdjGrads(r=gr.data_ptr(), dens=gd.data_ptr(), fdrag=gf.data_ptr())
    fl = sa._fields(x["C"], r=x["r"], dens=x["dens"], fdrag=x["fdrag"], shep=x["shep"])
    args = (int(x["q"].numel()), sa._ptr(x["q"]), sa._ptr(x["off"]), sa._ptr(x["b"]),
            sa._ptr(x["img"]), ctypes.byref(fl), ctypes.byref(x["C"].consts),
            sa._ptr(x["gpor"]), sa._ptr(x["gdrag"]))
    if noalloc:
        rc = lib.sphdem_adj_coupling_bwd_noalloc(*args, x["n"], ctypes.byref(g), sa._stream())
    else:
        work = torch.zeros_like(x["dens"])
        rc = lib.sphdem_adj_coupling_bwd(*args, sa._ptr(work), ctypes.byref(g), sa._stream())
    assert rc == 0
    torch.cuda.synchronize()
    return gr.cpu().numpy(), gd.cpu().numpy(), gf.cpu().numpy()


@pytest.mark.gpu
def test_noalloc_coupling_backward_equals_the_work_buffer_form():
    big = sd.adjoint_case(nf=12, nz=16, n_dem=200, seed=1007, vscale=0.1, touching=20, floor=3)
    small = sd.adjoint_case(nf=8, nz=12, n_dem=40, seed=1011, vscale=0.1)
    assert small.n < big.n
    for tag, case, seed in (("first call", big, 21), ("second call, smaller", small, 22)):
        print(f"noalloc coupling backward: {tag}, n={case.n}", flush=True)
        x = _coupling_inputs(case, seed)
        ref = _backward(x, noalloc=False)
        got = _backward(x, noalloc=True)
        for name, a, b in zip(("r", "dens", "fdrag"), got, ref):
            assert float(np.abs(b).max()) > 0, (tag, name)
            # the two forms differ only in the order of their FP64 atomics
            assert normwise(a, b) < 1e-12, (tag, name, normwise(a, b))
        x["pairs"].close()


def test_noalloc_coupling_backward_refuses_more_than_its_scratch():
    """Refused before any device work (no GPU needed)."""
    lib = sa._adj_lib()
    fl, cs, g = sa.AdjFields(), sa.AdjConsts(), sa.AdjGrads()
    buf = ctypes.c_void_p(0x1000)  # never dereferenced: the call is refused first
    too_many = (1 << 20) + 1
    assert lib.sphdem_adj_coupling_bwd_noalloc(
        3, buf, buf, buf, buf, ctypes.byref(fl), ctypes.byref(cs), buf, buf, too_many,
        ctypes.byref(g), None) == sd.E_CAPACITY
