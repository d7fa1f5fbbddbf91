#!/usr/bin/env python
"""Benchmark of the HiG-Flow projection time step on one MI355X.

Metric: cell-updates/s per full time step, mixed precision
versus the plain FP64 GPU step, plus the Poisson solve time per step.
``value`` is the mixed-precision step's cell-updates/s (cells x steps / wall
seconds of the timed steps, fields resident in HBM); the FP64 step, their
ratio, the Poisson milliseconds and the mixed-vs-FP64 errors are reported
beside it.  ``roofline`` is measured live for the dominant kernel with
hipEvents on the library stream; ``cpu_baseline`` times the CPU oracle (a
restatement of the reference algorithm) on the host on a bounded sample.

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--workload NAME]
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

PEAK_HBM_GBS = 8000.0  # MI355X HBM3E

# Workloads.  The default is the largest configuration that runs on one GPU:
# the batched sweep of 64 Oldroyd-B cavities (1.2 M cells).  The 2D
# Oldroyd-B planar 4:1 contraction on its 3-level HiG-Tree AMR mesh and the
# cavities are smaller cases of the same step.  `roof` names the kernel the
# roofline line reports: the one that dominates the workload's kernel trace,
# as (hfh_probe id, name, form probed).
WORKLOADS = {
    # channel of length 20, half-width 4 -> 1, base cells 0.1, 3 factor-2
    # levels within 1.0 of the re-entrant corners (47,800 leaves), parabolic
    # inflow of mean 1, Re=0.01, Wi=1, beta=1/9, log-conformation from the
    # identity, semi-implicit Euler, 200 steps.  dt=2e-5: at dt=1e-4 the
    # restated algorithm itself diverges (the CPU oracle at step 17).
    "contraction2d_3lev_oldroydb": dict(kind="contraction", levels=3, dt=2e-5, steps=200),
    # the same contraction on a finer, graded mesh: base cells 0.05 and 4
    # levels whose boxes shrink with the level (radius 1.0, 0.75, 0.5,
    # 0.25), so each level boundary is a 2:1 jump (121,000 leaves), dt=1e-5
    "contraction2d_4lev_graded": dict(kind="contraction", levels=4, graded=True, h0=0.05, dt=1e-5, steps=200),
    # 2D Newtonian lid-driven cavity, uniform 64x64, Re=100, dt=1e-3,
    # explicit Euler, incremental projection, CUBISTA, 100 steps
    "cavity2d_64_re100": dict(kind="cavity", n=64, Re=100.0, dt=1e-3, steps=100),
    # the same cavity on 256x256 (dt within the explicit viscous bound)
    "cavity2d_256_re100": dict(kind="cavity", n=256, Re=100.0, dt=1e-4, steps=100),
    # 2D Oldroyd-B cavity (log-conformation), uniform 128x128
    "ve_cavity2d_128": dict(kind="ve_cavity", n=128, Re=1.0, Wi=1.0, beta=0.5, dt=1e-5, steps=100),
    # one large simulation: Oldroyd-B cavity, uniform 1024x1024 (1,048,576
    # cells), Re=10, Wi=0.5, beta=0.5, semi-implicit Euler, dt=2e-4
    "ve_cavity2d_1024": dict(kind="ve_cavity", n=1024, Re=10.0, Wi=0.5, beta=0.5, dt=2e-4, steps=50, semi=True),
    # batched sweep: 64 independent Oldroyd-B cavities from the seed-0
    # PCG32 table (sides 32..256, Re log-uniform in [1, 100], Wi in
    # [0.1, 2]; 1,225,216 cells), beta=0.5, dt=5e-4, semi-implicit Euler,
    # all stepped together (one launch per stage)
    # 3D Newtonian lid-driven cavity on [0,1]^3: base n^3 cells and 2
    # factor-2 levels in a 0.05-thick layer along every wall (each level's
    # layer snapped outward to its parent grid), Re=400, lid u=1 on z=1,
    # dt=2e-4, explicit Euler, 100 steps, solves to 1e-9 (the stated
    # refinement target of this configuration); 16^3 base: 89,272 leaves,
    # 24^3 base: 310,120 leaves
    "cavity3d_16_wall2": dict(kind="wall", n=16, Re=400.0, dt=2e-4, steps=100, tol=1e-9,
                              roof=(15, "k_amg_resid<double> (pressure V-cycle, fine level, FP64 form)", "fp64")),
    "cavity3d_24_wall2": dict(kind="wall", n=24, Re=400.0, dt=2e-4, steps=100, tol=1e-9,
                              roof=(15, "k_amg_resid<double> (pressure V-cycle, fine level, FP64 form)", "fp64")),
    "sweep64_ve_cavity": dict(kind="sweep", count=64, seed=0, dt=5e-4, steps=100,
                              roof=(13, "k_amg_prolong_post<float> (velocity V-cycle, fine level, mixed form)",
                                    "mixed")),
}
DEFAULT = "sweep64_ve_cavity"
# launch A of the fused FP64 BiCGStab on the semi-implicit velocity block
# system (x, r, p updates and v = (A D^-1) p in one launch; the pressure
# system's on explicit workloads)
DEFAULT_ROOF = (9, "k_kbcg_a<double>", "fp64")


def make_case(spec, **tol):
    from higflow_amd import cases
    if spec["kind"] == "sweep":
        return cases.sweep(spec["count"], spec["seed"], dt=spec["dt"], **tol)
    if spec["kind"] == "cavity":
        return cases.cavity(spec["n"], Re=spec["Re"], dt=spec["dt"], **tol)
    if spec["kind"] == "wall":
        return cases.wall_layer_cavity(spec["n"], Re=spec["Re"], dt=spec["dt"], **tol)
    if spec["kind"] == "contraction":
        return cases.contraction(levels=spec["levels"], dt=spec["dt"], graded=spec.get("graded", False),
                                 h0=spec.get("h0", 0.1), **tol)
    if spec.get("semi"):
        tol = dict(tol, tempdiscrtype=cases.SEMI_IMPLICIT_EULER)
    return cases.ve_cavity(spec["n"], Re=spec["Re"], Wi=spec["Wi"], beta=spec["beta"], dt=spec["dt"], **tol)


def run_gpu(case, steps, warmup, **over):
    from higflow_amd import native
    sim = native.Batch(case, **over) if isinstance(case, list) else native.Flow(case, **over)
    for _ in range(warmup):
        sim.step()
    sim.sync()
    t0 = time.perf_counter()
    for _ in range(steps):
        sim.step()
    sim.sync()
    wall = time.perf_counter() - t0
    return sim, wall


def stage_pass(sim, steps):
    sim.reset_timing()
    sim.set_timing(True)
    for _ in range(steps):
        sim.step()
    sim.sync()
    sim.set_timing(False)
    return {k: v[0] / max(v[1], 1) for k, v in sim.timing().items()}


def probe(sim, which, reps):
    return sim.probe(which, reps)


def cpu_baseline(case, ncells, budget_s=15.0, max_steps=400):
    """The CPU oracle (FP64, BiCGStab + ILU(0), one core) on the first steps
    of the same workload, until `budget_s` seconds of step time have run
    (a sweep: its simulations in order, one step each).  Returns
    (cell-updates/s, steps, seconds)."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import hfo
    members = case if isinstance(case, list) else [case]
    done, cells_done, wall = 0, 0, 0.0
    sims = []
    for c in members:
        o = hfo.Oracle(c)
        t0 = time.perf_counter()
        o.step()
        wall += time.perf_counter() - t0
        sims.append(o)
        cells_done += o.size(1)
        done += 1
        if wall >= budget_s:
            break
    while len(members) == 1 and done < max_steps and wall < budget_s:
        t0 = time.perf_counter()
        sims[0].step()
        wall += time.perf_counter() - t0
        cells_done += ncells
        done += 1
    return cells_done / wall, done, wall


def _oracle_member_step(row):
    """Worker of the all-core leg: one oracle step of one sweep simulation
    (side, Re, Wi, dt, solver tolerances); returns (cells, step seconds)."""
    sys.path.insert(0, ROOT)
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import hfo
    from higflow_amd import cases
    n, re, wi, dt, tol = row
    o = hfo.Oracle(cases.sweep(table=[(n, re, wi)], dt=dt, **tol)[0])
    t0 = time.perf_counter()
    o.step()
    return o.size(1), time.perf_counter() - t0


def cpu_baseline_all_cores(rows, workers, budget_s=20.0):
    """The CPU oracle on every host core at once: the simulations of a
    sweep are independent, so `workers` processes each step whole
    simulations (one step each; `rows` as _oracle_member_step takes them)
    until the members whose steps have started within `budget_s` are done.
    Returns (cell-updates/s over the wall time of the stepping, simulations
    stepped, seconds), or None without rows (a single simulation: the
    oracle steps one mesh on one core)."""
    if not rows:
        return None
    import multiprocessing as mp
    ctx = mp.get_context("spawn")  # no GPU state in the workers
    cells = 0
    done = 0
    t0 = time.perf_counter()
    with ctx.Pool(workers) as pool:
        # completed steps count as they finish; work still running when the
        # budget ends is not counted, so the rate errs low
        for n, _ in pool.imap_unordered(_oracle_member_step, rows):
            cells += n
            done += 1
            if time.perf_counter() - t0 >= budget_s:
                break
        wall = time.perf_counter() - t0
        pool.terminate()
    return cells / wall, done, wall


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=None, help="timed steps (default: the workload's)")
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--workload", default=DEFAULT, choices=sorted(WORKLOADS))
    ap.add_argument("--no-cpu", action="store_true")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py: this build runs the projection step on one GPU; multi-GPU runs are not part of it "
              "(--gpus %d requested)" % args.gpus, file=sys.stderr)
        sys.exit(2)
    # set-up progress lines on stderr (3D meshes take minutes of host set-up)
    os.environ.setdefault("HFH_PROGRESS", "1")
    spec = WORKLOADS[args.workload]
    if args.steps is None:
        args.steps = spec["steps"]
    tol = spec.get("tol", 1e-10)
    ref_tol = dict(rtol=tol, atol=tol, maxit=50000)
    case = make_case(spec, **ref_tol)

    batch = isinstance(case, list)
    sim64, wall64 = run_gpu(case, args.steps, args.warmup, precision=0)
    views64 = [sim64.member(m) for m in range(sim64.n)] if batch else [sim64]
    ncells = sum(v.size(1) for v in views64)
    print("bench: %s, %d cells, FP64 %d steps in %.3f s" % (args.workload, ncells, args.steps, wall64),
          file=sys.stderr, flush=True)
    simmx, wallmx = run_gpu(case, args.steps, args.warmup, precision=1, inner_rtol=1e-5)
    viewsmx = [simmx.member(m) for m in range(simmx.n)] if batch else [simmx]
    c0 = case[0] if batch else case
    cups64 = ncells * args.steps / wall64
    cupsmx = ncells * args.steps / wallmx
    print("bench: mixed %d steps in %.3f s" % (args.steps, wallmx), file=sys.stderr, flush=True)
    errs = {}
    D = c0["dim"]
    fields = [("u%d" % d, 0, d) for d in range(D)] + [("p", 1, 0)]
    if c0["params"].get("flowtype") == 3:
        fields += [("psi%d%d" % (i, j), 6, i * D + j) for i in range(D) for j in range(i, D)]
    for name, f, c in fields:
        # a sweep: the worst simulation
        e = dict(l2=0.0, linf=0.0)
        for vm, v6 in zip(viewsmx, views64):
            a, b = vm.get(f, c), v6.get(f, c)
            nb = np.linalg.norm(b)
            e["l2"] = max(e["l2"], float(np.linalg.norm(a - b) / nb) if nb > 0 else 0.0)
            e["linf"] = max(e["linf"], float(np.abs(a - b).max() / max(np.abs(b).max(), 1e-300)))
        errs[name] = e
    k = max(5, min(20, args.steps // 5))
    st64 = stage_pass(sim64, k)
    stmx = stage_pass(simmx, k)
    conv, iters, resid, _ = sim64.last_solve()

    # the dominant kernel of the workload's kernel trace (its `roof` entry;
    # by default launch A of the fused FP64 BiCGStab on the semi-implicit
    # velocity block system, or the pressure system's for explicit runs),
    # timed live with hipEvents on the library stream
    roof = None
    rid, rname, rform = spec.get("roof", DEFAULT_ROOF)
    pr = probe(simmx if rform == "mixed" else sim64, rid, 400)
    if pr:
        ms, by = pr
        ach = by / (ms * 1e-3) / 1e9
        traffic = None
        tf = os.path.join(ROOT, "profiles", "traffic_%s.json" % args.workload)
        if os.path.exists(tf):
            traffic = json.load(open(tf)).get("roof_kernel_bytes_per_launch")
        roof = {"bound": "hbm", "achieved": round(ach, 2), "peak": PEAK_HBM_GBS, "unit": "GB/s",
                "frac": round(ach / PEAK_HBM_GBS, 5), "traffic": traffic,
                "kernel": rname, "bytes_per_launch": by, "avg_us": round(ms * 1e3, 3)}
    kernels = {}
    for which, name, sim in ((9, "k_kbcg_a<double>", sim64), (11, "k_kcg_spmv<double>(velocity)", sim64),
                             (12, "k_kcg_spmv<float>(velocity)", simmx),
                             (13, "k_amg_prolong_post<float>(velocity, fine level)", simmx),
                             (14, "k_amg_prolong_post<float>(pressure, fine level)", simmx),
                             (15, "k_amg_resid<double>(pressure, fine level)", sim64),
                             (10, "k_kbcg_b<double>", sim64), (6, "k_kbcg_b<double>(pressure)", sim64),
                             (7, "k_kbcg_b<float>(pressure)", simmx), (0, "k_spmv_dot<double>", sim64),
                             (1, "k_momentum", sim64), (2, "k_interp(momentum rows)", sim64),
                             (3, "k_interp(divergence rows)", sim64), (4, "k_interp(projection rows)", sim64),
                             (5, "k_constitutive", sim64), (8, "k_copy(64 MiB)", sim64)):
        p = probe(sim, which, 200)
        if p:
            kernels[name] = {"avg_us": round(p[0] * 1e3, 3), "bytes": p[1],
                             "GBps": round(p[1] / (p[0] * 1e-3) / 1e9, 2)}
    cpu = None
    if not args.no_cpu:
        v, n, wall = cpu_baseline(case, ncells)
        one = {"value": round(v, 1), "unit": "cell-updates/s", "cores": 1, "kind": "port",
               "sample": "%d %s of %s on the CPU oracle (BiCGStab+ILU(0), rtol=atol=%g), %.1f s"
                         % (n, "simulations stepped once" if batch else "steps", args.workload, tol, wall)}
        cpu = one
        workers = min(os.cpu_count() or 1, 16)
        rows = None
        if batch:
            from higflow_amd import cases
            rows = [(n, re, wi, spec["dt"], ref_tol) for (n, re, wi) in cases.sweep_table(spec["count"], spec["seed"])]
        allc = cpu_baseline_all_cores(rows, workers) if workers > 1 else None
        if allc:
            va, na, wa = allc
            cpu = {"value": round(va, 1), "unit": "cell-updates/s", "cores": workers, "kind": "port",
                   "sample": "%d simulations of %s stepped once each on %d worker processes of the CPU oracle "
                             "(BiCGStab+ILU(0), rtol=atol=%g), %.1f s" % (na, args.workload, workers, tol, wa),
                   "one_core": one}
    line = {
        "metric": "cell-updates/s per full time step (mixed vs FP64 GPU) + Poisson solve ms",
        "value": round(cupsmx, 1),
        "unit": "cell-updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(wallmx / args.steps * 1e3, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64-refinement/f32-inner",
        "data": "synthetic",
        "config": {"workload": args.workload, "cells": ncells, "dim": D,
                   "dt": c0["params"]["dt"],
                   "Re": [m["params"]["Re"] for m in case] if batch else c0["params"]["Re"],
                   "simulations": len(case) if batch else 1,
                   "flowtype": "viscoelastic" if c0["params"].get("flowtype") == 3 else "newtonian",
                   "time_scheme": "semi-implicit" if c0["params"].get("tempdiscrtype") == 3 else "explicit"},
        "fp64_value": round(cups64, 1),
        "fp64_ms_per_step": round(wall64 / args.steps * 1e3, 4),
        "mixed_over_fp64": round(cupsmx / cups64, 3),
        "poisson_ms_fp64": round(st64["solve"], 4),
        "poisson_ms_mixed": round(stmx["solve"], 4),
        "poisson_iters_fp64": iters,
        "err_mixed_vs_fp64": errs,
        "stage_ms_fp64": {k2: round(v2, 4) for k2, v2 in st64.items()},
        "stage_ms_mixed": {k2: round(v2, 4) for k2, v2 in stmx.items()},
        "solves_last_step": ({"fp64": [sim64.solve_stats(w) for w in range(D + 1)],
                              "mixed": [simmx.solve_stats(w) for w in range(D + 1)]} if not batch else
                             {"fp64_max_iters": [max(sim64.solve_stats(m, w)["iters"] for m in range(sim64.n))
                                                 for w in range(D + 1)],
                              "fp64_all_converged": all(sim64.solve_stats(m, w)["converged"] == 1
                                                        for m in range(sim64.n) for w in range(D + 1)),
                              # per system: the mixed form's last solve (0 FP32 refinement, 1 FP32
                              # failed and finished in FP64, 2 FP64 chosen), its iterations and
                              # corrections (the same for every member of a block-diagonal solve)
                              "mixed_form": [{k3: simmx.solve_stats(0, w)[k3] for k3 in ("fallback", "iters", "outer")}
                                             for w in range(D + 1)],
                              "mixed_all_converged": all(simmx.solve_stats(m, w)["converged"] == 1
                                                         for m in range(simmx.n) for w in range(D + 1))}),
        "roofline": roof,
        "kernels": kernels,
        "cpu_baseline": cpu,
    }
    print(json.dumps(line))


if __name__ == "__main__":
    main()
