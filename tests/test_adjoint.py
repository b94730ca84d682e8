"""The adjoint of the timestep loop (sphdem_adjoint): the differentiable step
against the oracle's steps, its gradient against central finite differences
of the oracle (CPU); the hand-written HIP backward kernels against torch
autograd of the same passes, and the HIP adjoint against finite differences of
the library's own forward (GPU)."""
import ctypes
import dataclasses
import os
import re

import numpy as np
import pytest
import torch

import oracle
import sphdem as sd
import sphdem_adjoint as sa
from adjoint_ref import OraclePairs, TorchPasses

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
FIELDS = ("r", "v", "vhat", "dens", "h", "f", "dddt")


def small_case():
    # 768 fluid, 384 wall, 10 DEM: one touching pair and one sphere on the floor
    return sd.adjoint_case(nf=8, nz=12, n_dem=10, seed=1004, vscale=0.1, touching=1, floor=1)


def normwise(a, b):
    a, b = np.asarray(a), np.asarray(b)
    return float(np.abs(a - b).max() / max(np.abs(b).max(), 1e-300))


def oracle_loss(case, steps, target, params=None, v=None):
    c = dataclasses.replace(case, v=case.v.copy(), vhat=case.vhat.copy())
    if v is not None:
        c.v[:] = v
        c.vhat[:] = v
    o = oracle.Oracle(c, params=params)
    o.run(steps)
    z = o.get("r")[case.iam == sd.DEM, 2]
    o.close()
    return float(np.mean((z - target) ** 2))


def test_adjoint_structs_mirror_the_header():
    text = open(os.path.join(ROOT, "include", "sphdem_adjoint.h")).read()
    for cls, name in ((sa.AdjConsts, "sphdem_adj_consts"), (sa.AdjFields, "sphdem_adj_fields"),
                      (sa.AdjGrads, "sphdem_adj_grads")):
        body = text[text.index("typedef struct " + name):text.index("} " + name + ";")]
        names = re.findall(r"\*?\s*(\w+)(?:\[\d+\])?;", body)
        assert [f[0] for f in cls._fields_] == names, name


def test_adjoint_entry_points_check_their_arguments():
    """Argument checks return before any device work (no GPU needed)."""
    lib = sa._adj_lib()
    fl, cs, g = sa.AdjFields(), sa.AdjConsts(), sa.AdjGrads()
    buf = ctypes.c_void_p(0x1000)  # never dereferenced: every call below is refused or empty
    for nm in ("omega", "dddt", "force", "fpi"):
        fn = getattr(lib, "sphdem_adj_" + nm)
        extra = [buf, buf] if nm == "force" else [buf]
        assert fn(-1, buf, buf, buf, buf, ctypes.byref(fl), ctypes.byref(cs), *extra, None) == -1
        assert fn(0, buf, buf, buf, buf, ctypes.byref(fl), ctypes.byref(cs), *extra, None) == 0
        # pairs given but the fields they read are missing
        assert fn(4, buf, buf, buf, buf, ctypes.byref(fl), ctypes.byref(cs), *extra, None) == -1
    assert lib.sphdem_adj_coupling(3, buf, buf, buf, buf, ctypes.byref(fl), ctypes.byref(cs),
                                   buf, buf, None, None) == -1
    assert lib.sphdem_adj_coupling_bwd(3, buf, buf, buf, buf, ctypes.byref(fl), ctypes.byref(cs),
                                       buf, buf, None, ctypes.byref(g), None) == -1
    assert lib.sphdem_adj_contacts(-2, buf, buf, buf, ctypes.byref(fl), ctypes.byref(cs), buf,
                                   None) == -1


def test_step_model_follows_the_oracle():
    case = small_case()
    model = sa.StepModel(case, device="cpu", passes=TorchPasses(), pairs=OraclePairs(case))
    orc = oracle.Oracle(case)
    th = model.params(requires_grad=False)
    S = model.initial_state()
    with torch.no_grad():
        for k in range(4):
            S = model.step(S, th, k)
            orc.run(1)
            assert float(S["dt"]) == pytest.approx(orc.globals().dt, rel=1e-13)
            for nm in FIELDS:
                assert normwise(model.original(S[nm]).numpy(), orc.get(nm)) < 1e-12, (k, nm)
            dem = case.iam == sd.DEM
            fd = model.original(S["fdrag"]).numpy()
            assert normwise(fd[dem], orc.get("fdrag")[dem]) < 1e-12


def test_gradient_matches_oracle_finite_differences():
    """Reverse mode through 6 steps (checkpoints every 2) against central
    differences of the oracle.  Parameter steps are 1e-3 relative: the loss
    moves by ~1e-19 at 1e-6, below its rounding (~1e-20 per DEM)."""
    case = small_case()
    steps = 6
    target = sd.adjoint_target(case)
    model = sa.StepModel(case, device="cpu", passes=TorchPasses(), pairs=OraclePairs(case))
    loss, g = model.gradient(steps, torch.as_tensor(target), every=2)
    assert loss == pytest.approx(oracle_loss(case, steps, target), rel=1e-12)
    tol = {"drag_a": 1e-4, "dem_k": 1e-4, "dem_gamma": 1e-4, "viscosity": 1e-4}
    for k, t in tol.items():
        base = getattr(case.params, k)
        eps = 1e-3 * abs(base)
        Pp, Pm = case.params.copy(), case.params.copy()
        setattr(Pp, k, base + eps)
        setattr(Pm, k, base - eps)
        fd = (oracle_loss(case, steps, target, Pp) - oracle_loss(case, steps, target, Pm)) / (2 * eps)
        assert g[k] == pytest.approx(fd, rel=t), k
    # the initial velocities: a random direction and single DEM components, step 1e-6
    d = sd.splitmix64(7, case.n * 3).reshape(case.n, 3) - 0.5
    d[case.iam == sd.SPH_BOUNDARY] = 0.0
    eps = 1e-6
    fd = (oracle_loss(case, steps, target, v=case.v + eps * d)
          - oracle_loss(case, steps, target, v=case.v - eps * d)) / (2 * eps)
    assert float((g["v_init"].numpy() * d).sum()) == pytest.approx(fd, rel=1e-4)
    for i in np.where(case.iam == sd.DEM)[0][:3]:
        e = np.zeros_like(d)
        e[i, 2] = 1.0
        fd = (oracle_loss(case, steps, target, v=case.v + eps * e)
              - oracle_loss(case, steps, target, v=case.v - eps * e)) / (2 * eps)
        assert g["v_init"][i, 2].item() == pytest.approx(fd, rel=1e-4)


def test_checkpoint_spacing_does_not_change_the_gradient():
    case = small_case()
    target = torch.as_tensor(sd.adjoint_target(case))
    model = sa.StepModel(case, device="cpu", passes=TorchPasses(), pairs=OraclePairs(case))
    ref = model.gradient(4, target, every=4)
    for every in (1, 3):
        got = model.gradient(4, target, every=every)
        assert got[0] == ref[0]
        for k in sa.PARAMS:
            assert got[1][k] == pytest.approx(ref[1][k], rel=1e-12, abs=1e-300), (every, k)
        assert normwise(got[1]["v_init"].numpy(), ref[1]["v_init"].numpy()) < 1e-12


def test_model_rejects_what_it_does_not_restate():
    case = small_case()
    c = dataclasses.replace(case, params=case.params.copy())
    c.params.reinit_every = 5
    with pytest.raises(ValueError, match="reinitialisation"):
        sa.StepModel(c, device="cpu", passes=TorchPasses(), pairs=None)
    with pytest.raises(ValueError, match="flag set"):
        sa.StepModel(sd.dambreak_case(nf=6), device="cpu", passes=TorchPasses(), pairs=None)


def test_hip_passes_refuse_cpu_tensors():
    """No CPU fallback: the HIP passes take GPU tensors only."""
    r = torch.zeros((4, 3), dtype=torch.float64)
    with pytest.raises(ValueError, match="GPU"):
        sa.HipPasses().contacts(None, None, r, r, None, None)


# ------------------------------------------------------------------ GPU
def _gpu_state(case, steps):
    """A state `steps` steps in, from the GPU model with torch passes."""
    pairs = sa.LibraryPairs(case)
    model = sa.StepModel(case, device="cuda", passes=TorchPasses(), pairs=pairs)
    th = model.params(requires_grad=False)
    S = model.initial_state()
    with torch.no_grad():
        S = model.rollout(S, th, steps)
    return model, pairs, S


@pytest.mark.gpu
def test_device_pair_export_equals_the_host_export():
    case = sd.adjoint_case(nf=12, nz=16, n_dem=200, seed=1007, vscale=0.1, touching=20, floor=3)
    pairs = sa.LibraryPairs(case)
    r = torch.as_tensor(case.r, device="cuda")
    h = torch.as_tensor(case.h, device="cuda")
    kinds = (sd.PAIRS_SPH, sd.PAIRS_DEM, sd.PAIRS_SPH2DEM, sd.PAIRS_DEM2SPH)
    got = pairs(r, h, kinds)  # first search: buffers sized from the counts
    for k in kinds:
        host = pairs.sim.pairs(k)
        dev = np.stack([c.cpu().numpy() for c in got[k]], axis=1)
        assert len(host) > 0 and np.array_equal(dev, host), k
    # a second search at moved positions reuses (or regrows) those buffers
    gen = torch.Generator(device="cuda").manual_seed(3)
    r2 = r + 0.2 * float(case.params.dem_radius) * (
        torch.rand(r.shape, generator=gen, device="cuda", dtype=torch.float64) - 0.5)
    got = pairs(r2, h, kinds)
    for k in kinds:
        host = pairs.sim.pairs(k)
        dev = np.stack([c.cpu().numpy() for c in got[k]], axis=1)
        assert len(host) > 0 and np.array_equal(dev, host), k
    pairs.close()


@pytest.mark.gpu
def test_set_field_device_equals_set_field():
    case = sd.adjoint_case(nf=12, nz=16, n_dem=200, seed=1007)
    a, b = sd.Sim(case), sd.Sim(case)
    rng = np.random.default_rng(8)
    for nm, shape in (("r", (case.n, 3)), ("h", (case.n,)), ("vhat", (case.n, 3))):
        x = np.asarray(getattr(case, nm), float) + 1e-4 * rng.standard_normal(shape)
        a.set(nm, x)
        t = torch.as_tensor(x, device="cuda").contiguous()
        torch.cuda.synchronize()
        b.set_device(nm, t.data_ptr())
        assert np.array_equal(a.get(nm), b.get(nm)), nm
        assert np.array_equal(b.get(nm), x), nm
    with pytest.raises(sd.SphDemError):
        t = torch.zeros(case.n, dtype=torch.int32, device="cuda")
        b.set_device("iam", t.data_ptr())
    a.close()
    b.close()


@pytest.mark.gpu
def test_hip_passes_match_torch_autograd():
    case = sd.adjoint_case(nf=12, nz=16, n_dem=200, seed=1007, vscale=0.1, touching=20, floor=3)
    print(f"adjoint passes: {case.counts()}", flush=True)
    model, pairs, S = _gpu_state(case, 2)
    th = model.params(requires_grad=True)
    C = sa._Ctx(model, th)
    r, h = S["r"], S["h"]
    ps = model._pairs(("passes", 0), r, h,
                      (sd.PAIRS_SPH, sd.PAIRS_SPH2DEM, sd.PAIRS_DEM2SPH, sd.PAIRS_DEM))
    assert ps[sd.PAIRS_DEM].count > 0
    gen = torch.Generator(device="cuda").manual_seed(5)
    rnd = lambda t: torch.rand(t.shape, generator=gen, dtype=t.dtype, device=t.device) + 0.5
    dens = S["dens"]
    omega = TorchPasses().omega(C, ps[sd.PAIRS_SPH], r, h, dens).detach()
    pdr2 = torch.where(model.grp, 1e-3 * rnd(h), torch.zeros_like(h))
    por = torch.where(model.grp, 0.5 + 0.4 * torch.rand(h.shape, generator=gen, dtype=h.dtype,
                                                          device=h.device), torch.ones_like(h))
    fp = torch.where(model.grp[:, None], rnd(r) - 1.0, torch.zeros_like(r))
    fv = torch.where(model.grp[:, None], rnd(r) - 1.0, torch.zeros_like(r))
    fdrag = torch.where(model.dem[:, None], rnd(r) - 1.0, torch.zeros_like(r))
    cases = {
        "omega": (lambda P, x: P.omega(C, ps[sd.PAIRS_SPH], *x), [r, h, dens], []),
        "dddt": (lambda P, x: P.dddt(C, ps[sd.PAIRS_SPH], *x), [r, S["vhat"], h, omega], []),
        "coupling": (lambda P, x: P.coupling(C, ps[sd.PAIRS_SPH2DEM], *x), [r, dens, fdrag], []),
        "force": (lambda P, x: P.force(C, ps[sd.PAIRS_SPH], *x[:6], th["viscosity"]),
                  [r, S["vhat"], h, dens, omega, pdr2], ["viscosity"]),
        "fpi": (lambda P, x: P.fpi(C, ps[sd.PAIRS_DEM2SPH], *x, th["viscosity"], th["drag_a"],
                                   th["drag_b"]),
                [r, S["vhat"], h, dens, fp, fv, por], ["viscosity", "drag_a", "drag_b"]),
        "contacts": (lambda P, x: P.contacts(C, ps[sd.PAIRS_DEM], *x, th["dem_k"],
                                             th["dem_gamma"]),
                     [r, S["vhat"]], ["dem_k", "dem_gamma"]),
    }
    for name, (fn, xs, pnames) in cases.items():
        outs = {}
        for tag, P in (("hip", sa.HipPasses()), ("torch", TorchPasses())):
            x = [t.detach().clone().contiguous().requires_grad_(True) for t in xs]
            for k in pnames:
                th[k].grad = None
            y = fn(P, x)
            y = y if isinstance(y, tuple) else (y,)
            gy = [torch.sin(7.0 * torch.arange(t.numel(), device=t.device,
                                               dtype=t.dtype)).reshape(t.shape) for t in y]
            torch.autograd.backward(y, gy)
            outs[tag] = ([t.detach() for t in y], [t.grad for t in x],
                         [float(th[k].grad) for k in pnames])
        (yh, gh, ph), (yt, gt, pt) = outs["hip"], outs["torch"]
        for a, b in zip(yh, yt):
            assert float(b.abs().max()) > 0, name
            assert normwise(a.cpu(), b.cpu()) < 1e-12, name
        for i, (a, b) in enumerate(zip(gh, gt)):
            assert normwise(a.cpu(), b.cpu()) < 1e-10, (name, i)
        for k, a, b in zip(pnames, ph, pt):
            assert a == pytest.approx(b, rel=1e-9, abs=1e-300), (name, k)
        print(f"  {name}: forward and backward agree", flush=True)
    pairs.close()


def production_loss(case, steps, target, params=None, v=None):
    c = dataclasses.replace(case, v=case.v.copy(), vhat=case.vhat.copy(),
                            params=(params or case.params).copy())
    if v is not None:
        c.v[:] = v
        c.vhat[:] = v
    sim = sd.Sim(c)
    sim.run(steps)
    z = sim.get("r")[case.iam == sd.DEM, 2]
    sim.close()
    return float(np.mean((z - target) ** 2))


@pytest.mark.gpu
def test_hip_adjoint_matches_finite_differences_of_the_library():
    case = small_case()
    steps = 6
    target = sd.adjoint_target(case)
    pairs = sa.LibraryPairs(case)
    model = sa.StepModel(case, device="cuda", pairs=pairs)
    print("adjoint: small case, 6 steps, checkpoints every 2", flush=True)
    loss, g = model.gradient(steps, torch.as_tensor(target, device="cuda"), every=2)
    assert loss == pytest.approx(production_loss(case, steps, target), rel=1e-11)
    for k in ("drag_a", "dem_k", "dem_gamma", "viscosity"):
        base = getattr(case.params, k)
        eps = 1e-3 * abs(base)
        Pp, Pm = case.params.copy(), case.params.copy()
        setattr(Pp, k, base + eps)
        setattr(Pm, k, base - eps)
        fd = (production_loss(case, steps, target, Pp)
              - production_loss(case, steps, target, Pm)) / (2 * eps)
        assert g[k] == pytest.approx(fd, rel=1e-4), k
    d = sd.splitmix64(7, case.n * 3).reshape(case.n, 3) - 0.5
    d[case.iam == sd.SPH_BOUNDARY] = 0.0
    eps = 1e-6
    fd = (production_loss(case, steps, target, v=case.v + eps * d)
          - production_loss(case, steps, target, v=case.v - eps * d)) / (2 * eps)
    assert float((g["v_init"].cpu().numpy() * d).sum()) == pytest.approx(fd, rel=1e-4)
    # the torch expressions of the same passes give the same adjoint
    model_t = sa.StepModel(case, device="cuda", passes=TorchPasses(), pairs=pairs)
    loss_t, g_t = model_t.gradient(steps, torch.as_tensor(target, device="cuda"), every=2)
    assert loss_t == pytest.approx(loss, rel=1e-12)
    for k in sa.PARAMS:
        assert g[k] == pytest.approx(g_t[k], rel=1e-8, abs=1e-300), k
    assert normwise(g["v_init"].cpu(), g_t["v_init"].cpu()) < 1e-9
    pairs.close()


def test_gathered_backward_refuses_an_asymmetric_pair_set():
    """The backward kernels of force, omega and dddt gather the partner terms
    from the query's own list: a set missing a mirror pair (b, a, 26 - image)
    is refused before any launch, and a symmetric one passes the check."""
    L = torch.tensor([1.0, 1.0, 0.0], dtype=torch.float64)
    sym = sa.PairSet(torch.tensor([[0, 1, 13], [1, 0, 13], [2, 3, 12], [3, 2, 14]]), 4, L)
    sym.require_symmetric()
    lop = sa.PairSet(torch.tensor([[0, 1, 13], [1, 0, 13], [2, 3, 12], [3, 2, 12]]), 4, L)
    with pytest.raises(ValueError, match="not symmetric"):
        lop.require_symmetric()
    r = torch.zeros((4, 3), dtype=torch.float64)
    h = torch.ones(4, dtype=torch.float64)
    for call in (lambda P: P.omega(None, lop, r, h, h), lambda P: P.dddt(None, lop, r, r, h, h),
                 lambda P: P.force(None, lop, r, r, h, h, h, h, None)):
        with pytest.raises(ValueError, match="not symmetric"):
            call(sa.HipPasses())
