"""Benchmark: SPH-DEM particle-steps/s of the gfx950 timestep loop.

Workload (one GPU): SPH-DEM sedimentation, 64^3 = 262,144 fluid particles
(dx = 1/64 m, periodic in x and y), 2 x 3 layers of sphBoundary walls
(24,576) and 4,096 DEM spheres (diameter dx/2, density 2500), the sediment3D
flag set, FP64.  A step is one run.cpp iteration (sphStep++, the sediment3D
hooks, start/middle/end) over all particles with the inputs already resident
in HBM.  particle-steps/s = (N_fluid + N_DEM) * steps / wall time of the
timed steps (walls are processed but not counted).

Alongside the metric the JSON line carries
  roofline      : the dominant kernel's algorithmic bytes per launch divided by
                  its mean duration, timed with HIP events on the library's own
                  stream over the timed steps, against the 8 TB/s HBM peak
                  (default kernel: the SPH neighbour-list build, two launches a
                  step); `traffic` is the kernel's HBM bytes per launch from the
                  rocprofv3 FETCH_SIZE / WRITE_SIZE summary committed under
                  profiles/, scaled by the copy-kernel calibration it records;
  step_roofline : whole-step algorithmic bytes (1152 B per SPH/wall particle +
                  650 B per DEM particle per step) x steps/s against 8 TB/s;
  cpu_baseline  : the CPU FP64 restatement timed on this host on a bounded
                  sample of the same workload.
"""
import argparse
import json
import math
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "sph-dem_amd"))

import numpy as np  # noqa: E402

import sphdem as sd  # noqa: E402

METRIC = "SPH-DEM particle-steps/s, mixed vs FP64 at stated error; % HBM roofline"
HBM_PEAK_GBS = 8000.0

# Algorithmic bytes per particle per launch (FP64 widths), from the pass table:
# each input field read once, each output field written once per pass.
B_ALG_KERNEL = {
    # calcNeighbours: r h -> the list (4 B per entry, filled in below from the
    # measured mean list length) + its count
    "lists": 32 + 4,
    # calcForce with calcOmega in the same walk (the step's omega pass, VAR_H):
    # r vhat h mass dens pdr2 iam porosity -> f fp fv fb omega
    "force": 76 + 16 + 24 + 72 + 8,
    # calcDddtDudt with calcOmega in the same walk: r vhat h mass dens iam ->
    # dddt omega
    "dddt": 76 + 8 + 8,
    # porosity (28 -> 8) + DEM->SPH drag (56 -> 48)
    "coupling": 36 + 104,
}
B_ALG_STEP_SPH = 1152  # per SPH / wall particle per step (SPH-DEM flag set)
B_ALG_STEP_DEM = 650   # per DEM particle per step


def parse():
    p = argparse.ArgumentParser()
    p.add_argument("--gpus", type=int, default=1)
    p.add_argument("--steps", type=int, default=200)
    p.add_argument("--warmup", type=int, default=10)
    p.add_argument("--workload", choices=("sediment", "large", "ensemble", "adjoint"),
                   default="sediment",
                   help="sediment: 64^3 sedimentation (the headline); large: the 160^3 "
                        "dam-break geometry with 65,536 DEM and the SPH-DEM flag set; "
                        "ensemble: 256 independent sedimentation boxes of 4,096-65,536 "
                        "fluid particles advanced together, against the same boxes run "
                        "one after another")
    p.add_argument("--boxes", type=int, default=256, help="ensemble: number of boxes")
    p.add_argument("--no-tune", action="store_true",
                   help="ensemble: keep the default tile shape (no autotuning)")
    p.add_argument("--workers", type=int, default=8,
                   help="ensemble: host threads stepping the boxes")
    p.add_argument("--precision", choices=("fp64", "mixed"), default=None,
                   help="ensemble: arithmetic form (default mixed)")
    p.add_argument("--nf", type=int, default=None, help="fluid particles per axis")
    p.add_argument("--ndem", type=int, default=None)
    p.add_argument("--roofline-kernel", default="lists")
    p.add_argument("--traffic-file", default=None,
                   help="rocprofv3 FETCH_SIZE/WRITE_SIZE summary to take `traffic` from "
                        "(default: the committed summary of the default workload)")
    p.add_argument("--cpu-seconds", type=float, default=12.0,
                   help="budget of the CPU baseline sample")
    p.add_argument("--no-cpu", action="store_true")
    p.add_argument("--no-mixed", action="store_true",
                   help="skip the mixed-precision form (its timing and error)")
    p.add_argument("--copy-calibration", action="store_true",
                   help="only run a plain 512 MiB device-to-device copy (three times): the "
                        "reference kernel that scales rocprofv3's FETCH_SIZE / WRITE_SIZE")
    return p.parse_args()


def copy_calibration():
    import torch

    n = 64 << 20  # doubles: 512 MiB
    src = torch.ones(n, dtype=torch.float64, device="cuda")
    dst = torch.empty_like(src)
    for _ in range(3):
        dst.copy_(src)
    torch.cuda.synchronize()
    print(json.dumps({"copy_calibration_bytes": n * 8, "copies": 3}), flush=True)


def cpu_baseline(case, budget_s):
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import oracle

    orc = oracle.Oracle(case)
    done = 0
    t0 = time.perf_counter()
    while True:
        orc.run(1)
        done += 1
        el = time.perf_counter() - t0
        if el >= budget_s or done >= 5:
            break
    counted = int(np.sum((case.iam == sd.SPH) | (case.iam == sd.DEM)))
    orc.close()
    return {
        "value": counted * done / el,
        "unit": "particle-steps/s",
        "cores": 1,
        "kind": "port",
        "sample": f"{done} step(s) of the same {case.n}-particle case "
                  f"({el:.1f} s, single-threaded CPU FP64 restatement)",
    }


KERNEL_SYMBOL = {"lists": "k_lists", "force": "k_force_l",
                 "dddt": "k_dddt_l", "coupling": "k_coupling_sph",
                 "adj_force_bwd": "k_force_bwd"}


def traffic_from_summary(path, kernel):
    """HBM bytes per launch of `kernel` from a committed rocprofv3 summary:
    lines `calibration FETCH_SIZE <factor>` / `calibration WRITE_SIZE <factor>`
    and `<symbol> <dispatches> <FETCH_SIZE KiB> <WRITE_SIZE KiB>` per kernel."""
    if not os.path.exists(path):
        return None
    fac = {}
    rows = {}
    with open(path) as f:
        for line in f:
            w = line.split()
            if not w or w[0].startswith("#"):
                continue
            if w[0] == "calibration" and len(w) >= 3:
                fac[w[1]] = float(w[2])
            elif len(w) >= 4 and w[1].isdigit():
                rows[w[0]] = (float(w[2]), float(w[3]))
    sym = KERNEL_SYMBOL.get(kernel)
    if sym not in rows or "FETCH_SIZE" not in fac or "WRITE_SIZE" not in fac:
        return None
    fetch, write = rows[sym]
    return 1024.0 * (fetch * fac["FETCH_SIZE"] + write * fac["WRITE_SIZE"])


def make_case(a):
    if a.workload == "large":
        nf = a.nf or 160
        return sd.dambreak_case(nf=nf, dx=0.32 / nf, seed=1002,
                                with_dem=a.ndem if a.ndem is not None else 65536,
                                liq_dem_flags=True)
    nf = a.nf or 64
    return sd.sediment_case(nf=nf, n_dem=a.ndem if a.ndem is not None else 4096, seed=1001)


def workload_name(a, counts):
    if a.workload == "large":
        return (f"SPH-DEM dam-break geometry {round(counts['fluid'] ** (1 / 3))}^3 fluid + "
                f"{counts['wall']} wall + {counts['dem']} DEM, closed box, SPH-DEM flags, FP64")
    return (f"SPH-DEM sedimentation {round(counts['fluid'] ** (1 / 3))}^3 fluid + "
            f"{counts['wall']} wall + {counts['dem']} DEM, periodic x/y, sediment3D flags, FP64")


def _errors(mix, ref, iam):
    sph, dem = iam == sd.SPH, iam == sd.DEM
    e = {"dens_rel_max": float(np.max(np.abs(mix["dens"][sph] - ref["dens"][sph])
                                      / np.abs(ref["dens"][sph])))}
    if "f" in ref:
        fs = ref["f"][sph]
        e["f_sph_normwise"] = float(np.abs(mix["f"][sph] - fs).max() / np.abs(fs).max())
        if dem.any():
            fd = ref["f"][dem]
            e["f_dem_normwise"] = float(np.abs(mix["f"][dem] - fd).max() / np.abs(fd).max())
    if dem.any():
        e["dem_r_max_abs"] = float(np.abs(mix["r"][dem] - ref["r"][dem]).max())
    return e


def mixed_form(a, counts, counted, final64):
    """The mixed-precision form (FP32 pair arithmetic over FP64 positions, lists,
    sums, integration and dt) on the same workload: its throughput over the same
    timed steps, and its error against the FP64 form after one step (arithmetic
    error) and after warmup + timed steps (including the growth of differences
    along the trajectory)."""
    one = {}
    for prec in (sd.FP64, sd.MIXED):
        s = sd.Sim(make_case(a), device=0, precision=prec)
        s.run(1)
        one[prec] = {k: s.get(k) for k in ("r", "dens", "f")}
        iam = s.get("iam")
        s.close()
    sim = sd.Sim(make_case(a), device=0, precision=sd.MIXED)
    sim.run(a.warmup)
    sim.sync()
    t0 = time.perf_counter()
    sim.run(a.steps)
    sim.sync()
    el = time.perf_counter() - t0
    fin = {k: sim.get(k) for k in ("r", "dens")}
    sim.close()
    return {
        "value": counted * a.steps / el,
        "unit": "particle-steps/s",
        "ms_per_step": 1e3 * el / a.steps,
        "dtype": "f32 pair arithmetic, f64 positions/lists/sums/integration/dt",
        "error_vs_fp64_1_step": _errors(one[sd.MIXED], one[sd.FP64], iam),
        "error_vs_fp64_after_run": _errors(fin, final64, iam),
        "steps_compared_after_run": a.warmup + a.steps,
    }


def ensemble(a):
    """Independent boxes (sd.ensemble_cases) advanced together by
    sphdem_run_steps_many, then the same boxes one after another (sphdem_run_steps
    per box), both timed over `steps` steps after `warmup` steps together.
    particle-steps/s counts fluid + DEM of every box."""
    prec = sd.FP64 if a.precision == "fp64" else sd.MIXED
    cases = sd.ensemble_cases(nbox=a.boxes, seed=1005)
    sims = [sd.Sim(c, device=0, precision=prec) for c in cases]
    counted = sum(int(np.sum((c.iam == sd.SPH) | (c.iam == sd.DEM))) for c in cases)
    fluid = sum(int(np.sum(c.iam == sd.SPH)) for c in cases)
    dem = sum(int(np.sum(c.iam == sd.DEM)) for c in cases)
    wall = sum(int(np.sum(c.iam == sd.SPH_BOUNDARY)) for c in cases)
    sd.run_many(sims, a.warmup, workers=a.workers)
    for s in sims:
        s.sync()
    untuned = None
    shapes = {}
    if not a.no_tune:
        # the default shape first, then the autotuner: one representative box
        # per power-of-two fluid-count bucket is tuned (sphdem_tune_tile_shape,
        # state restored) and its shape given to the bucket
        t0 = time.perf_counter()
        sd.run_many(sims, a.steps, workers=a.workers)
        for s in sims:
            s.sync()
        untuned = time.perf_counter() - t0
        buckets = {}
        for s, c in zip(sims, cases):
            buckets.setdefault(int(math.log2(int(np.sum(c.iam == sd.SPH)))), []).append(s)
        for k, group in sorted(buckets.items()):
            seg, blk, ms = group[0].tune_tile_shape(steps=3)
            for s in group:
                s.set_tile_shape(seg, blk)
            shapes[f"fluid 2^{k}"] = {"boxes": len(group), "seg_particles": seg,
                                      "walker_block": blk, "tuned_ms_per_step": ms}
    t0 = time.perf_counter()
    sd.run_many(sims, a.steps, workers=a.workers)
    for s in sims:
        s.sync()
    el = time.perf_counter() - t0
    t1 = time.perf_counter()
    for s in sims:
        s.run(a.steps)
        s.sync()
    seq = time.perf_counter() - t1
    for s in sims:
        if not np.all(np.isfinite(s.get("dens"))):
            raise RuntimeError("non-finite density after the run")
        s.close()
    step_bytes = B_ALG_STEP_SPH * (fluid + wall) + B_ALG_STEP_DEM * dem
    step_gbs = step_bytes * (a.steps / el) / 1e9
    out = {
        "metric": METRIC,
        "value": counted * a.steps / el,
        "unit": "particle-steps/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": 1e3 * el / a.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64" if prec == sd.FP64 else
                 "f32 pair arithmetic, f64 positions/lists/sums/integration/dt",
        "data": "synthetic",
        "config": {
            "workload": f"ensemble of {a.boxes} sedimentation boxes, fluid 4096-65536 per box "
                        "(log-uniform), DEM = fluid/16, drag/stiffness/damping x U(0.8, 1.2)",
            "boxes": a.boxes, "fluid": fluid, "wall": wall, "dem": dem, "seed": 1005,
            "workers": a.workers,
        },
        "sequential": {
            "value": counted * a.steps / seq,
            "unit": "particle-steps/s",
            "ms_per_step": 1e3 * seq / a.steps,
            "what": "the same boxes one after another (sphdem_run_steps per box)",
        },
        "speedup_vs_sequential": seq / el,
        "tile_shapes": shapes or None,
        "default_shape": None if untuned is None else {
            "value": counted * a.steps / untuned, "unit": "particle-steps/s",
            "what": "the same boxes together with the default tile shape (64 particles per "
                    "segment, 256-thread walker blocks), timed before tuning"},
        "roofline": None,
        "step_roofline": {
            "alg_bytes_per_step": step_bytes,
            "achieved": step_gbs,
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": step_gbs / HBM_PEAK_GBS,
        },
    }
    print(json.dumps(out), flush=True)


# Algorithmic bytes of the backward SPH force kernel: per SPH-group particle
# its inputs r vhat h dens mass omega pdr2 iam and output adjoints gfv gfp read
# once (140 B) and its input adjoints r vhat h dens omega pdr2 written once
# (80 B); per pair the (a, b, image) triple (12 B).
B_ADJ_FORCE_BWD_PARTICLE = 220
B_ADJ_PAIR = 12


def adjoint(a):
    """Gradient of a 32-step rollout (sd.adjoint_case: 64 x 64 x 16 fluid,
    8,192 DEM) by the checkpointed reverse mode of sphdem_adjoint: the loss is
    the mean squared DEM z against sd.adjoint_target; gradients with respect
    to the drag coefficients, contact stiffness and damping, viscosity and the
    initial velocities.  value = (fluid + DEM) * rollout steps / wall time of
    one gradient (forward rollout + recomputation + backward)."""
    import torch
    import sphdem_adjoint as sa

    steps = a.steps if a.steps != 200 else 32
    nf = a.nf or 64
    ndem = a.ndem or 8192
    case = sd.adjoint_case(nf=nf, nz=16, n_dem=ndem)
    counts = case.counts()
    counted = counts["fluid"] + counts["dem"]
    target_np = sd.adjoint_target(case)
    target = torch.as_tensor(target_np, device="cuda")
    pairs = sa.LibraryPairs(case)
    model = sa.StepModel(case, device="cuda", pairs=pairs)
    model.gradient(min(2, steps), target, every=8)                   # warm-up
    torch.cuda.synchronize()
    pairs.seconds = 0.0
    sa.KERNEL_TIMES = {}
    t0 = time.perf_counter()
    loss, g = model.gradient(steps, target, every=8)
    torch.cuda.synchronize()
    el = time.perf_counter() - t0
    kt = sa.kernel_ms()
    sa.KERNEL_TIMES = None
    search_s = pairs.seconds
    peak_mem = torch.cuda.max_memory_allocated() / 2 ** 30
    # the library's own forward over the same rollout, and finite differences of it
    def prod_loss(P=None, v=None):
        import dataclasses
        c = dataclasses.replace(case, v=case.v.copy(), vhat=case.vhat.copy(),
                                params=(P or case.params).copy())
        if v is not None:
            c.v[:] = v
            c.vhat[:] = v
        sim = sd.Sim(c)
        sim.sync()
        t = time.perf_counter()
        sim.run(steps)
        sim.sync()
        dt = time.perf_counter() - t
        z = sim.get("r")[case.iam == sd.DEM, 2]
        sim.close()
        return float(np.mean((z - target_np) ** 2)), dt
    base, fwd_s = prod_loss()
    fd = {}
    for k in ("drag_a", "dem_k", "viscosity"):
        b0 = getattr(case.params, k)
        row = {"adjoint": g[k]}
        for rel in (1e-6, 1e-3):
            Pp, Pm = case.params.copy(), case.params.copy()
            setattr(Pp, k, b0 * (1 + rel))
            setattr(Pm, k, b0 * (1 - rel))
            d = (prod_loss(Pp)[0] - prod_loss(Pm)[0]) / (2 * rel * abs(b0))
            row[f"fd_rel_step_{rel:g}"] = d
            row[f"rel_err_{rel:g}"] = abs(g[k] - d) / abs(d) if d else None
        fd[k] = row
    dirn = sd.splitmix64(7, case.n * 3).reshape(case.n, 3) - 0.5
    dirn[case.iam == sd.SPH_BOUNDARY] = 0.0
    adj_dir = float((g["v_init"].cpu().numpy() * dirn).sum())
    eps = 1e-6
    d = (prod_loss(v=case.v + eps * dirn)[0] - prod_loss(v=case.v - eps * dirn)[0]) / (2 * eps)
    fd["v_init_direction"] = {"adjoint": adj_dir, "fd_step_1e-06": d,
                              "rel_err_1e-06": abs(adj_dir - d) / abs(d) if d else None,
                              "what": "all fluid and DEM velocities"}
    dd = np.where((case.iam == sd.DEM)[:, None], dirn, 0.0)
    adj_dd = float((g["v_init"].cpu().numpy() * dd).sum())
    d = (prod_loss(v=case.v + eps * dd)[0] - prod_loss(v=case.v - eps * dd)[0]) / (2 * eps)
    fd["v_init_dem_direction"] = {"adjoint": adj_dd, "fd_step_1e-06": d,
                                  "rel_err_1e-06": abs(adj_dd - d) / abs(d) if d else None,
                                  "what": "DEM velocities only"}
    kernels = {k: {"ms_per_launch": ms / n, "launches": n, "items_per_launch": it / n}
               for k, (ms, n, it) in kt.items()}
    roof = None
    if "force_bwd" in kt:
        ms, n, it = kt["force_bwd"]
        units = counts["fluid"] + counts["wall"]
        b_launch = B_ADJ_FORCE_BWD_PARTICLE * units + B_ADJ_PAIR * it / n
        ach = b_launch / (ms / n / 1e3) / 1e9
        tfile = os.path.join(ROOT, "profiles", "traffic_adjoint_64x64x16.txt")
        traffic = (traffic_from_summary(tfile, "adj_force_bwd")
                   if nf == 64 and ndem == 8192 else None)
        roof = {"bound": "hbm", "kernel": "adj_force_bwd", "achieved": ach, "peak": HBM_PEAK_GBS,
                "unit": "GB/s", "frac": ach / HBM_PEAK_GBS, "traffic": traffic,
                "mean_launch_ms": ms / n, "launches": n, "alg_bytes_per_launch": b_launch}
    out = {
        "metric": METRIC,
        "value": counted * steps / el,
        "unit": "particle-steps/s",
        "n_gpus": 1,
        "steps": steps,
        "warmup": 1,
        "ms_per_step": 1e3 * el / steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {
            "workload": f"adjoint: sedimentation {nf}x{nf}x16 fluid + {ndem} DEM, {steps}-step "
                        "rollout, loss = mean squared DEM z vs a seeded target, gradient wrt "
                        "drag coefficients, contact stiffness/damping, viscosity and initial "
                        "velocities, state checkpoint every 8 steps",
            "fluid": counts["fluid"], "wall": counts["wall"], "dem": counts["dem"], "seed": 1004,
        },
        "loss": loss,
        "loss_library_forward": base,
        "library_forward_particle_steps_per_s": counted * steps / fwd_s,
        "gradient_over_forward_time": el / fwd_s,
        "gradient_seconds": el,
        "pair_search_seconds": search_s,
        "peak_device_memory_gib": peak_mem,
        "grads": {k: g[k] for k in sa.PARAMS},
        "fd_check": fd,
        "kernels": kernels,
        "roofline": roof,
    }
    pairs.close()
    if not a.no_cpu:
        out["cpu_baseline"] = adjoint_cpu_baseline()
    print(json.dumps(out), flush=True)


def adjoint_cpu_baseline():
    """The same differentiable step on the host (torch on the CPU, the torch
    expressions of the pair passes, the oracle's pair searches): the gradient
    of a 1-step rollout of a 32 x 32 x 16 box of the same geometry (the same
    neighbour counts and DEM fraction per particle), timed."""
    import torch
    import sphdem_adjoint as sa
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    from adjoint_ref import OraclePairs, TorchPasses

    case = sd.adjoint_case(nf=32, nz=16, n_dem=2048)
    target_np = sd.adjoint_target(case)
    threads = torch.get_num_threads()
    pairs = OraclePairs(case)
    model = sa.StepModel(case, device="cpu", passes=TorchPasses(), pairs=pairs)
    t0 = time.perf_counter()
    model.gradient(1, torch.as_tensor(target_np), every=8)
    el = time.perf_counter() - t0
    pairs.close()
    counted = int(np.sum((case.iam == sd.SPH) | (case.iam == sd.DEM)))
    return {"value": counted / el, "unit": "particle-steps/s", "cores": threads, "kind": "port",
            "sample": "gradient of a 1-step rollout of a 32x32x16 sedimentation box with 2048 "
                      "DEM on the host (torch CPU model, oracle pair searches)"}


def main():
    a = parse()
    if a.copy_calibration:
        copy_calibration()
        return
    if a.gpus != 1:
        print(f"bench.py: this build runs on one GPU; --gpus {a.gpus} is not supported "
              "(the path shards by space, see the design notes)", file=sys.stderr)
        sys.exit(2)
    if a.workload == "ensemble":
        ensemble(a)
        return
    if a.workload == "adjoint":
        adjoint(a)
        return
    case = make_case(a)
    counts = case.counts()
    counted = counts["fluid"] + counts["dem"]
    sim = sd.Sim(case, device=0)
    sim.run(a.warmup)
    sim.sync()
    sim.timing_begin(a.roofline_kernel)
    t0 = time.perf_counter()
    sim.run(a.steps)
    sim.sync()
    t1 = time.perf_counter()
    kms, launches = sim.timing_end()
    el = t1 - t0
    # sanity: the state is finite
    for nm in ("r", "dens"):
        if not np.all(np.isfinite(sim.get(nm))):
            raise RuntimeError(f"non-finite {nm} after the run")
    g = sim.globals()
    units = counts["fluid"] + counts["wall"]  # SPH-group particles a pair kernel processes
    mean_nb = sim.pair_count(sd.PAIRS_SPH) / max(units, 1)
    final64 = {k: sim.get(k) for k in ("r", "dens")}
    sim.close()
    mixed = None if a.no_mixed else mixed_form(a, counts, counted, final64)
    value = counted * a.steps / el
    if a.roofline_kernel in ("fpi", "contacts", "shepsum"):
        units = counts["dem"]
    per_unit = B_ALG_KERNEL.get(a.roofline_kernel, 0)
    if a.roofline_kernel == "lists":
        per_unit += 4 * mean_nb
    b_launch = per_unit * units
    tfile = a.traffic_file
    if tfile is None and a.workload == "sediment" and a.nf is None and a.ndem is None:
        tfile = os.path.join(ROOT, "profiles", "traffic_sediment_64cubed.txt")
    traffic = traffic_from_summary(tfile, a.roofline_kernel) if tfile else None
    mean_s = (kms / 1e3) / max(launches, 1)
    achieved = b_launch / mean_s / 1e9 if launches else None
    step_bytes = B_ALG_STEP_SPH * (counts["fluid"] + counts["wall"]) + B_ALG_STEP_DEM * counts["dem"]
    step_gbs = step_bytes * (a.steps / el) / 1e9
    out = {
        "metric": METRIC,
        "value": value,
        "unit": "particle-steps/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": 1e3 * el / a.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {
            "workload": workload_name(a, counts),
            "fluid": counts["fluid"], "wall": counts["wall"], "dem": counts["dem"],
            "seed": 1002 if a.workload == "large" else 1001, "sim_time_s": g.time,
            "dt_last": g.dt,
        },
        "roofline": {
            "bound": "hbm",
            "kernel": a.roofline_kernel,
            "achieved": achieved,
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": (achieved / HBM_PEAK_GBS) if achieved else None,
            "traffic": traffic,
            "mean_launch_ms": mean_s * 1e3,
            "launches": launches,
            "alg_bytes_per_launch": b_launch,
            "mean_sph_neighbours": mean_nb,
        },
        "step_roofline": {
            "alg_bytes_per_step": step_bytes,
            "achieved": step_gbs,
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": step_gbs / HBM_PEAK_GBS,
        },
    }
    if mixed is not None:
        out["mixed"] = mixed
    if not a.no_cpu:
        out["cpu_baseline"] = cpu_baseline(case, a.cpu_seconds)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
