"""Benchmark: ns/cell per RK3 step and Poisson-solve GB/s on one MI355X.

Workload: the periodic channel of 512 x 256 x 192 cells (Lx = 4 pi, Ly = 2 pi,
Lz = 1, tanh-stretched walls with factor 1.8, Re_b = 5600, constant flow rate),
FP64, dt = 2e-4, synthetic seeded initial state (laminar profile + 10 % noise).
A step is one full RK3 step (3 substeps) through the fused C-ABI call; inputs
are resident in HBM.  The Poisson solve is timed separately with HIP events on
the context stream (B_P = 80 N bytes per FP64 solve: five passes, 8 B in + 8 B
out per cell each).  The oracle (C/OpenMP restatement) is timed on the host on
a bounded number of full steps of the same case as the CPU baseline.

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--ng N1 N2 N3] [--no-cpu]
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "ns/cell per RK3 step + Poisson GB/s (% HBM peak), mixed vs FP64, 1 GPU"
HBM_PEAK_GBS = 8000.0


def cpu_baseline(p, budget_s=20.0):
    """Oracle steps on the host (rank 0 only): ns/cell per RK3 step."""
    import numpy as np
    from tests import cases
    from oracle import oracle as O
    orc = cases.oracle_case(p)

    class G:
        pass

    g = G()
    g.dl, g.l, g.zc, g.dzf = orc.dl, orc.l, orc.zc, orc.dzf
    u, v, w, pr = cases.initial_state("channel", p, g, seed=1)
    st = orc.new_state()
    for k, q in zip("uvwp", (u, v, w, pr)):
        st[k][1:-1, 1:-1, 1:-1] = q
    orc.init_bounds(st)
    orc.rk3_step(st, 2e-4)  # warm-up (page-in, FFT tables)
    times = []
    t_all = time.perf_counter()
    while True:
        t0 = time.perf_counter()
        orc.rk3_step(st, 2e-4)
        times.append(time.perf_counter() - t0)
        if time.perf_counter() - t_all > budget_s or len(times) >= 10:
            break
    n = int(np.prod(p["ng"]))
    med = float(np.median(times))
    return {"value": med / n * 1e9, "unit": "ns/cell/step", "cores": O.lib().orc_num_threads(),
            "kind": "port",
            "sample": "%d full RK3 steps of the %dx%dx%d channel (median), oracle C/OpenMP, FP64"
                      % (len(times), *p["ng"])}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--ng", type=int, nargs=3, default=(512, 256, 192))
    ap.add_argument("--solves", type=int, default=10)
    ap.add_argument("--no-cpu", action="store_true")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py: this build runs on one GPU; multi-GPU runs (--gpus %d) are not part of it" % args.gpus,
              file=sys.stderr)
        sys.exit(2)

    import numpy as np
    import torch
    from cans_amd.host import Simulation
    from cans_amd import initflow as IF
    from tests import cases

    p = cases.params("channel", ng=tuple(args.ng))
    sim = Simulation(cases.host_case(p))
    u, v, w, pr = cases.initial_state("channel", p, sim.grid, seed=1)
    sim.set_state(*(IF.halo(p["ng"], q) for q in (u, v, w, pr)))
    sim.init_bounds()
    dt = 2e-4
    ncell = int(np.prod(p["ng"]))
    for _ in range(args.warmup):
        sim.rk3_step(dt)
    torch.cuda.synchronize()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    t0 = time.perf_counter()
    e0.record()
    for _ in range(args.steps):
        sim.rk3_step(dt)
    e1.record()
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    ms_step = e0.elapsed_time(e1) / args.steps
    # Poisson solve alone (same stream), repeated on the last right-hand side
    s0, s1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    sim.solver()
    torch.cuda.synchronize()
    s0.record()
    for _ in range(args.solves):
        sim.solver()
    s1.record()
    torch.cuda.synchronize()
    t_p = s0.elapsed_time(s1) / args.solves * 1e-3
    b_p = 80.0 * ncell
    gbs = b_p / t_p / 1e9
    # device copy of a known size: the box's streaming bandwidth (and the reference
    # point that scales the FETCH_SIZE / WRITE_SIZE counters of the kernels)
    nbytes = 1 << 30
    src = torch.ones(nbytes // 8, dtype=torch.float64, device="cuda")
    dst = torch.empty_like(src)
    dst.copy_(src)
    torch.cuda.synchronize()
    c0, c1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    c0.record()
    for _ in range(5):
        dst.copy_(src)
    c1.record()
    torch.cuda.synchronize()
    copy_gbs = 2 * nbytes * 5 / (c0.elapsed_time(c1) * 1e-3) / 1e9
    del src, dst
    tot, divmax = sim.chkdiv()
    out = {
        "metric": METRIC,
        "value": ms_step * 1e-3 / ncell * 1e9,
        "unit": "ns/cell/step",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_step,
        "higher_is_better": False,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (seeded laminar profile + 10% noise)",
        "config": {"workload": "periodic channel %dx%dx%d, stretched z (gr=1.8), Re_b=5600, dt=2e-4, FP64"
                               % tuple(p["ng"]), "cells": ncell, "precision": "fp64"},
        "poisson": {"ms": t_p * 1e3, "GBps": gbs, "pct_hbm_peak": 100.0 * gbs / HBM_PEAK_GBS,
                    "algorithmic_bytes": b_p},
        "roofline": {"bound": "hbm", "achieved": gbs, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": gbs / HBM_PEAK_GBS, "traffic": None,
                     "kernel": "Poisson solve (x-fwd, y-fwd, z-Thomas, y-bwd, x-bwd passes)"},
        "copy_GBps": copy_gbs,
        "divmax": divmax,
        "wall_s": wall,
    }
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline(p)
    print(json.dumps(out))


if __name__ == "__main__":
    main()
