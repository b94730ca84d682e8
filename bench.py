"""Benchmark: SPH-DEM particle-steps/s of the gfx950 timestep loop.

Default workload (one GPU): the large mixed-precision SPH-DEM case, 160^3 =
4,096,000 fluid particles (dx = 0.002 m) in the dam-break box (2L x L x L,
L = 0.32 m) closed by 3 layers of sphBoundary walls (791,256), with 65,536
DEM spheres (diameter dx/2, density 2500) in the upper half of the fluid
block, the SPH-DEM flag set (Morris viscosity, VAR_H_CORRECTION, LIQ_DEM,
Di Felice drag, custom DEM floor).  The same initial state is run in both
forms, FP64 and mixed (FP32 pair arithmetic over FP64 positions, lists,
sums, integration and dt); `value` is the mixed form's throughput and the
line carries the FP64 form's, their ratio and the mixed form's error against
FP64.  A step is one run.cpp iteration (sphStep++, the sediment3D hooks,
start/middle/end) over all particles with the inputs already resident in HBM.
particle-steps/s = (N_fluid + N_DEM) * steps / wall time of the timed steps
(walls are processed but not counted).  `--workload sediment` runs the 64^3
sedimentation box (periodic x/y) the same way.

Alongside the metric the JSON line carries
  roofline      : the dominant kernel of the headline form (default: the SPH
                  neighbour-list build, two launches a step) with its
                  algorithmic bytes per launch -- the cell-build row of the
                  per-pass table, 32 B per SPH/wall particle (read r, write
                  cell key and index) -- divided by its mean duration, timed
                  with HIP events on the library's own stream over the timed
                  steps, against the 8 TB/s HBM peak; `traffic` is the kernel's
                  HBM bytes per launch from the rocprofv3 FETCH_SIZE /
                  WRITE_SIZE summary committed under profiles/, scaled by the
                  copy-kernel calibration it records; the list entries the
                  kernel writes (4 B each) are reported beside it, not in it;
  step_roofline : whole-step algorithmic bytes of the headline form (FP64:
                  1152 B per SPH/wall particle + 650 B per DEM particle per
                  step; mixed: 1040 + 650) x steps/s against 8 TB/s;
  cpu_baseline  : the CPU FP64 restatement timed on this host's cores on one
                  step of the same workload (all host threads, OpenMP).
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

# Algorithmic bytes per SPH/wall particle per launch, from the per-pass table
# (each input field of a pass read once, each output field written once, at
# the widths the form reads and stores them; neighbour re-reads not counted).
# A kernel that does the work of several passes counts each of them.
B_ALG_KERNEL = {
    # the cell build (read r, write cell key + index): the list build's row
    "lists": {"fp64": 32, "mixed": 32},
    # calcForce (r vhat h mass dens pdr2 iam, omega porosity -> f fp fv fb)
    # with calcOmega summed in the same walk (r h mass dens iam -> omega)
    "force": {"fp64": 196 + 60, "mixed": 156 + 40},
    # calcDddtDudt (r vhat h mass iam, omega -> dddt) with calcOmega
    "dddt": {"fp64": 84 + 60, "mixed": 52 + 40},
    # porosity (28 -> 8) + DEM->SPH drag (56 -> 48)
    "coupling": {"fp64": 36 + 104, "mixed": 36 + 104},
}
# whole step per SPH/wall particle and per DEM particle
B_ALG_STEP_SPH = {"fp64": 1152, "mixed": 1040}
B_ALG_STEP_DEM = {"fp64": 650, "mixed": 650}


def parse():
    p = argparse.ArgumentParser()
    p.add_argument("--gpus", type=int, default=1)
    p.add_argument("--steps", type=int, default=None,
                   help="timed steps (default 100; the adjoint's rollout 32)")
    p.add_argument("--warmup", type=int, default=None,
                   help="untimed steps before the timed ones (default 5; the ensemble 40: "
                        "its first tens of steps in a process run slower)")
    p.add_argument("--workload", choices=("sediment", "large", "ensemble", "adjoint"),
                   default="large",
                   help="large (the headline): 160^3 fluid + 65,536 DEM in the closed "
                        "dam-break box with the SPH-DEM flag set, FP64 and mixed forms; "
                        "sediment: the 64^3 sedimentation box; ensemble: 256 independent "
                        "sedimentation boxes of 4,096-65,536 fluid particles advanced "
                        "together, against the same boxes run one after another")
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
    p.add_argument("--cpu-steps", type=int, default=1,
                   help="steps of the CPU baseline sample")
    p.add_argument("--cpu-threads", type=int, default=0,
                   help="host threads of the CPU baseline (0: all)")
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


def cpu_baseline(case, steps, threads):
    """The CPU FP64 restatement (oracle/, the checker of the GPU library) on
    this host: `steps` steps of the same case, its gather passes on `threads`
    OpenMP threads (0: all the host offers).  Only the steps are timed."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import oracle

    used = oracle.set_threads(threads if threads > 0 else (os.cpu_count() or 1))
    orc = oracle.Oracle(case)
    t0 = time.perf_counter()
    orc.run(steps)
    el = time.perf_counter() - t0
    counted = int(np.sum((case.iam == sd.SPH) | (case.iam == sd.DEM)))
    orc.close()
    return {
        "value": counted * steps / el,
        "unit": "particle-steps/s",
        "cores": used,
        "kind": "port",
        "sample": f"{steps} step(s) of the same {case.n}-particle case ({el:.1f} s, the CPU "
                  f"FP64 restatement, gather passes on {used} OpenMP threads, the resets "
                  "and the Shepard scatter serial)",
    }


KERNEL_SYMBOL = {"lists": "k_lists_t", "force": "k_force_l",
                 "dddt": "k_dddt_l", "coupling": "k_coupling_sph",
                 "adj_force_bwd": "k_force_bwd"}
KERNEL_SYMBOL_MIXED = {"force": "k_force_m", "dddt": "k_dddt_m"}


def traffic_from_summary(path, kernel, form="fp64"):
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
    if form == "mixed":
        sym = KERNEL_SYMBOL_MIXED.get(kernel, sym)
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
        return (f"SPH-DEM closed dam-break box {round(counts['fluid'] ** (1 / 3))}^3 fluid + "
                f"{counts['wall']} wall + {counts['dem']} DEM, SPH-DEM flags")
    return (f"SPH-DEM sedimentation {round(counts['fluid'] ** (1 / 3))}^3 fluid + "
            f"{counts['wall']} wall + {counts['dem']} DEM, periodic x/y, sediment3D flags")


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


def time_form(a, case, prec, kernel):
    """One form on the workload: the first warm-up step (its state is the
    one-step error reference), the other warm-up steps, then `steps` timed
    steps with every launch of `kernel` bracketed by HIP events on the
    library's stream."""
    sim = sd.Sim(case, device=0, precision=prec)
    one = None
    if a.warmup > 0:
        sim.run(1)
        one = {k: sim.get(k) for k in ("r", "dens", "f")}
        sim.run(a.warmup - 1)
    sim.sync()
    sim.timing_begin(kernel)
    t0 = time.perf_counter()
    sim.run(a.steps)
    sim.sync()
    el = time.perf_counter() - t0
    kms, launches = sim.timing_end()
    for nm in ("r", "dens"):
        if not np.all(np.isfinite(sim.get(nm))):
            raise RuntimeError(f"non-finite {nm} after the run")
    counts = case.counts()
    out = {
        "seconds": el,
        "kernel_ms": kms,
        "launches": launches,
        "one_step": one,
        "final": {k: sim.get(k) for k in ("r", "dens")},
        "iam": sim.get("iam"),
        "globals": sim.globals(),
        "mean_nb": sim.pair_count(sd.PAIRS_SPH) / max(counts["fluid"] + counts["wall"], 1),
    }
    sim.close()
    return out


def run_forms(a):
    """The FP64 and mixed forms on the same initial state; the mixed form is
    the headline unless --no-mixed."""
    case = make_case(a)
    counts = case.counts()
    counted = counts["fluid"] + counts["dem"]
    units = counts["fluid"] + counts["wall"]  # SPH-group particles a pair kernel processes
    kernel = a.roofline_kernel
    print(f"bench: {workload_name(a, counts)}; FP64 form, {a.warmup} + {a.steps} steps",
          file=sys.stderr, flush=True)
    f64 = time_form(a, case, sd.FP64, kernel)
    mix = None
    if not a.no_mixed:
        print("bench: mixed form", file=sys.stderr, flush=True)
        mix = time_form(a, case, sd.MIXED, kernel)
    head, form = (mix, "mixed") if mix is not None else (f64, "fp64")

    def rate(r):
        return counted * a.steps / r["seconds"]

    if kernel in ("fpi", "contacts", "shepsum"):
        units = counts["dem"]
    per_unit = B_ALG_KERNEL.get(kernel, {}).get(form, 0)
    b_launch = per_unit * units
    mean_s = (head["kernel_ms"] / 1e3) / max(head["launches"], 1)
    achieved = b_launch / mean_s / 1e9 if head["launches"] else None
    tfile = a.traffic_file
    if tfile is None and a.nf is None and a.ndem is None:
        tfile = os.path.join(ROOT, "profiles", {
            "large": "traffic_closedbox_160cubed.txt",
            "sediment": "traffic_sediment_64cubed.txt"}[a.workload])
    traffic = traffic_from_summary(tfile, kernel, form) if tfile else None
    step_bytes = B_ALG_STEP_SPH[form] * units + B_ALG_STEP_DEM[form] * counts["dem"]
    step_gbs = step_bytes * (a.steps / head["seconds"]) / 1e9
    g = head["globals"]
    out = {
        "metric": METRIC,
        "value": rate(head),
        "unit": "particle-steps/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": 1e3 * head["seconds"] / a.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32/f64 mixed" if form == "mixed" else "f64",
        "data": "synthetic",
        "config": {
            "workload": workload_name(a, counts),
            "form": form,
            "fluid": counts["fluid"], "wall": counts["wall"], "dem": counts["dem"],
            "seed": 1002 if a.workload == "large" else 1001, "sim_time_s": g.time,
            "dt_last": g.dt,
        },
        "roofline": {
            "bound": "hbm",
            "kernel": kernel,
            "achieved": achieved,
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": (achieved / HBM_PEAK_GBS) if achieved else None,
            "traffic": traffic,
            "mean_launch_ms": mean_s * 1e3,
            "launches": head["launches"],
            "alg_bytes_per_launch": b_launch,
            "alg_bytes_per_unit": per_unit,
            "units_per_launch": units,
            "mean_sph_neighbours": head["mean_nb"],
            "list_entry_bytes_per_launch": (4.0 * head["mean_nb"] * units
                                            if kernel == "lists" else None),
        },
        "step_roofline": {
            "alg_bytes_per_step": step_bytes,
            "achieved": step_gbs,
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": step_gbs / HBM_PEAK_GBS,
        },
    }
    f64_bytes = B_ALG_STEP_SPH["fp64"] * units + B_ALG_STEP_DEM["fp64"] * counts["dem"]
    out["fp64"] = {
        "value": rate(f64),
        "unit": "particle-steps/s",
        "ms_per_step": 1e3 * f64["seconds"] / a.steps,
        "step_roofline_frac": f64_bytes * (a.steps / f64["seconds"]) / 1e9 / HBM_PEAK_GBS,
        "kernel_mean_launch_ms": (f64["kernel_ms"] / max(f64["launches"], 1)),
    }
    if mix is not None:
        out["mixed_over_fp64"] = f64["seconds"] / mix["seconds"]
        err = {"after_run": _errors(mix["final"], f64["final"], f64["iam"]),
               "steps_compared_after_run": a.warmup + a.steps}
        if mix["one_step"] is not None:
            err["after_1_step"] = _errors(mix["one_step"], f64["one_step"], f64["iam"])
        out["mixed_error_vs_fp64"] = err
    if not a.no_cpu:
        print("bench: CPU baseline", file=sys.stderr, flush=True)
        out["cpu_baseline"] = cpu_baseline(case, a.cpu_steps, a.cpu_threads)
    print(json.dumps(out), flush=True)


def ensemble(a):
    """Independent boxes (sd.ensemble_cases) advanced together by
    sphdem_run_steps_many (every step of a box launched with its reset on the
    device and nothing read back until the box's last step), then the same
    boxes replayed from captured HIP graphs (sphdem_run_steps_many_graph), then
    one after another (sphdem_run_steps per box), each timed over `steps` steps
    after `warmup` steps together.  particle-steps/s counts fluid + DEM of
    every box."""
    prec = sd.FP64 if a.precision == "fp64" else sd.MIXED
    cases = sd.ensemble_cases(nbox=a.boxes, seed=1005)
    sims = [sd.Sim(c, device=0, precision=prec) for c in cases]
    counted = sum(int(np.sum((c.iam == sd.SPH) | (c.iam == sd.DEM))) for c in cases)
    fluid = sum(int(np.sum(c.iam == sd.SPH)) for c in cases)
    dem = sum(int(np.sum(c.iam == sd.DEM)) for c in cases)
    wall = sum(int(np.sum(c.iam == sd.SPH_BOUNDARY)) for c in cases)

    def together(steps, graph=False):
        t0 = time.perf_counter()
        sd.run_many(sims, steps, workers=a.workers, graph=graph)
        for s in sims:
            s.sync()
        return time.perf_counter() - t0

    together(a.warmup)  # each box's first step runs as launches with host resets
    untuned = None
    shapes = {}
    if not a.no_tune:
        # the default shape first, then the autotuner: one representative box
        # per power-of-two fluid-count bucket is tuned (sphdem_tune_tile_shape,
        # state restored) and its shape given to the bucket
        untuned = together(a.steps)
        buckets = {}
        for s, c in zip(sims, cases):
            buckets.setdefault(int(math.log2(int(np.sum(c.iam == sd.SPH)))), []).append(s)
        for k, group in sorted(buckets.items()):
            seg, blk, ms = group[0].tune_tile_shape(steps=3)
            for s in group:
                s.set_tile_shape(seg, blk)
            shapes[f"fluid 2^{k}"] = {"boxes": len(group), "seg_particles": seg,
                                      "walker_block": blk, "tuned_ms_per_step": ms}
        together(1)  # the first steps with the tuned shapes
    el = together(a.steps)
    together(1, graph=True)  # one step as launches, then each box's capture
    graphed = together(a.steps, graph=True)
    together(1, graph="ensemble")  # the graph over all boxes is built
    one_graph = together(a.steps, graph="ensemble")
    t1 = time.perf_counter()
    for s in sims:
        s.run(a.steps)
        s.sync()
    seq = time.perf_counter() - t1
    for s in sims:
        if not np.all(np.isfinite(s.get("dens"))):
            raise RuntimeError("non-finite density after the run")
        s.close()
    fm = "fp64" if prec == sd.FP64 else "mixed"
    step_bytes = B_ALG_STEP_SPH[fm] * (fluid + wall) + B_ALG_STEP_DEM[fm] * dem
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
        "graph": {
            "value": counted * a.steps / graphed, "unit": "particle-steps/s",
            "ms_per_step": 1e3 * graphed / a.steps,
            "what": "the same boxes together, each box's step replayed from its captured "
                    "HIP graph (sphdem_run_steps_many_graph)"},
        "ensemble_graph": {
            "value": counted * a.steps / one_graph, "unit": "particle-steps/s",
            "ms_per_step": 1e3 * one_graph / a.steps,
            "what": "the same boxes together, one graph holding every box's captured step "
                    "launched once per ensemble step (sphdem_run_steps_ensemble_graph)"},
        "tile_shapes": shapes or None,
        "default_shape": None if untuned is None else {
            "value": counted * a.steps / untuned, "unit": "particle-steps/s",
            "what": "the same boxes together with the default tile shape (per-target list "
                    "search, 256-thread walker blocks), timed before tuning"},
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

    steps = a.steps
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
    # the library's own forward over the same rollout, and central finite
    # differences of it.  Each tracked run steps through the phase calls with
    # the sediment3D hooks (as sphdem_run_steps does, bit for bit) and keeps,
    # after every reset, the sorted pair keys of each search kind; a perturbed
    # run then reports how many resets saw a pair set other than the base
    # run's.  Where none did, the loss is a smooth function of the perturbation
    # and the FD must agree with the adjoint to 1e-4.
    kinds = (sd.PAIRS_SPH, sd.PAIRS_DEM, sd.PAIRS_SPH2DEM, sd.PAIRS_DEM2SPH)

    class _Exporter:
        _pairs = sa.LibraryPairs._pairs

        def __init__(self, sim):
            self.sim, self.cap = sim, {}

    def snapshot(exp):
        out = []
        for k in kinds:
            pa, pb, pi = exp._pairs(k, "cuda")
            key = (pa.long() * case.n + pb.long()) * 27 + pi.long()
            out.append(torch.sort(key).values)
        return out

    def prod_loss(P=None, v=None, track=False):
        import dataclasses
        c = dataclasses.replace(case, v=case.v.copy(), vhat=case.vhat.copy(),
                                params=(P or case.params).copy())
        if v is not None:
            c.v[:] = v
            c.vhat[:] = v
        sim = sd.Sim(c)
        sim.sync()
        t = time.perf_counter()
        keys = None
        if not track:
            sim.run(steps)
        else:
            keys = []
            exp = _Exporter(sim)
            G = np.array([0.0, 0.0, sa.GRAVITY])
            td = c.params.time_drop
            for _ in range(steps):
                gl = sim.globals()
                gl.sphStep += 1
                sim.set_globals(gl)
                sim.start()
                sim.hook_body_force(1 << sd.SPH, G)
                sim.middle()
                keys.append(snapshot(exp))
                sim.hook_sum_drag()
                sim.hook_body_force(1 << sd.DEM, G)
                sim.hook_dem_floor()
                sim.hook_freeze(1 << sd.DEM, td)
                sim.end()
                keys.append(snapshot(exp))
                sim.hook_freeze(1 << sd.DEM, td)
        sim.sync()
        dt = time.perf_counter() - t
        z = sim.get("r")[case.iam == sd.DEM, 2]
        sim.close()
        return float(np.mean((z - target_np) ** 2)), dt, keys

    base, fwd_s, _ = prod_loss()
    base_t, _, base_keys = prod_loss(track=True)

    def changed(keys):
        """Resets (of 2 per step) where some pair set differs from the base run's."""
        return sum(any(not torch.equal(x, y) for x, y in zip(kb, kn))
                   for kb, kn in zip(base_keys, keys))

    def changed_by_kind(keys):
        """The same count for each pair kind alone."""
        return {name: sum(not torch.equal(kb[i], kn[i]) for kb, kn in zip(base_keys, keys))
                for i, name in enumerate(("sph", "dem", "sph2dem", "dem2sph"))}

    def central(run_p, run_m, denom, adj):
        lp, _, kp = run_p()
        lm, _, km = run_m()
        d = (lp - lm) / denom
        cp, cm = changed(kp), changed(km)
        err = abs(adj - d) / abs(d) if d else None
        bp, bm = changed_by_kind(kp), changed_by_kind(km)
        return {"fd": d, "rel_err": err, "resets_with_other_pairs": [cp, cm],
                "by_kind": {k: [bp[k], bm[k]] for k in bp},
                "pair_sets_identical": cp == 0 and cm == 0,
                "within_1e-4": bool(err is not None and err <= 1e-4)}

    fd = {"tracked_base_loss_equals_library_forward": base_t == base,
          "resets_per_run": 2 * steps}
    for k in ("drag_a", "dem_k", "viscosity"):
        b0 = getattr(case.params, k)
        row = {"adjoint": g[k]}
        for rel in (1e-6, 1e-3):
            Pp, Pm = case.params.copy(), case.params.copy()
            setattr(Pp, k, b0 * (1 + rel))
            setattr(Pm, k, b0 * (1 - rel))
            row[f"rel_step_{rel:g}"] = central(lambda: prod_loss(Pp, track=True),
                                               lambda: prod_loss(Pm, track=True),
                                               2 * rel * abs(b0), g[k])
        fd[k] = row
    dirn = sd.splitmix64(7, case.n * 3).reshape(case.n, 3) - 0.5
    dirn[case.iam == sd.SPH_BOUNDARY] = 0.0
    vin = g["v_init"].cpu().numpy()
    for name, sel, what in (("v_init_direction", case.iam != sd.SPH_BOUNDARY,
                             "all fluid and DEM velocities"),
                            ("v_init_fluid_direction", case.iam == sd.SPH,
                             "fluid velocities only"),
                            ("v_init_dem_direction", case.iam == sd.DEM, "DEM velocities only")):
        dd = np.where(sel[:, None], dirn, 0.0)
        adj = float((vin * dd).sum())
        row = {"adjoint": adj, "what": what}
        for eps in (1e-6, 1e-5, 1e-4):
            row[f"step_{eps:g}"] = central(lambda: prod_loss(v=case.v + eps * dd, track=True),
                                           lambda: prod_loss(v=case.v - eps * dd, track=True),
                                           2 * eps, adj)
        fd[name] = row
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
    if a.steps is None:
        a.steps = 32 if a.workload == "adjoint" else 100
    if a.warmup is None:
        a.warmup = 40 if a.workload == "ensemble" else 5
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
    run_forms(a)


if __name__ == "__main__":
    main()
