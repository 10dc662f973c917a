"""Benchmark: ns/cell per RK3 step and Poisson-solve GB/s on one MI355X, mixed vs FP64.

Workload: the periodic channel of 1024 x 512 x 384 cells (Lx = 4 pi, Ly = 2 pi,
Lz = 1, tanh-stretched walls with factor 1.8, Re_b = 5600, constant flow rate,
dt = 2e-4), synthetic seeded initial state (laminar profile + 10 % noise), run
twice from the same state: in the mixed mode (velocity stored as FP64; p, pp,
RHS_old, the RK increment and the transform passes stored as FP32; FFT stages
computing in FP32, every other kernel in FP64) and in plain FP64.  A step is one full RK3 step (3
substeps) through the fused C-ABI call; inputs are resident in HBM.  `value` is
the mixed mode's ns/cell/step, with two FP64 refinement sweeps of the last
substep's pressure solve on every 10th step (the steps whose divergence the
reference checks, icheck = 10; the last timed step is one of them, so the
reported divmax is that check's).  Beside it: the mixed step refined on every
step, the plain mixed step (no refinement), the FP64 step, the speed-ups and the
relative L-inf difference of the final fields (mixed vs FP64).

The Poisson solve is timed on its own with HIP events on the context stream:
B_P = 5 passes x N cells x (s_in + s_out) = 40 N bytes with FP32 storage, 80 N
with FP64 storage.  The oracle (C/OpenMP restatement) is timed on the host on a
bounded number of full steps of the same case as the CPU baseline.

"library_fft_reference" times rocFFT (through torch.fft) on the Poisson solve's x and y
transforms of the same field, for scale beside the solve's own transform passes.

Secondary workloads, timed after the headline one and reported under "secondary":
the 256^3 heated cube with a scalar, buoyancy and 3-D implicit diffusion (FP64 and mixed),
64 independent Poisson boxes of different sizes solved as one batch (FP64 and mixed
storage, beside the same boxes solved one at a time), and the gradient of the kinetic
energy after 32 steps of the 256^3 Taylor-Green flow and after 16 steps of a 512x256x192
channel with walls in z (forward and adjoint, FP64).

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--ng N1 N2 N3]
                       [--precision both|mixed|fp64] [--solves S] [--no-cpu] [--no-secondary]
                       [--cavity-only] [--boxes-only] [--adjoint-only [tgv|channel]]
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
DT = 2e-4
ICHECK = 10  # steps between divergence checks (icheck of every reference example input)
# HBM bytes moved by one standalone Poisson solve, measured with rocprofv3 --pmc
# (FETCH_SIZE x 2 + WRITE_SIZE x 1, the factors a 1 GiB device copy calibrates);
# the per-kernel table is profiles/channel_1024x512x384_hbm_traffic.md
SOLVE_TRAFFIC_BYTES = {((1024, 512, 384), "mixed"): 9.184e9, ((1024, 512, 384), "fp64"): 18.045e9}
TRAFFIC_SOURCE = "profiles/channel_1024x512x384_hbm_traffic.md"


def cpu_baseline(p, init, budget_s=20.0):
    """Oracle steps on the host (rank 0 only): ns/cell per RK3 step."""
    import numpy as np
    from tests import cases
    from oracle import oracle as O
    try:  # built with -march=native on this host, as the reference's CPU build would be
        march = O.use_native()
    except Exception as e:  # noqa: BLE001 - fall back to the portable build, and say so
        print("[cpu] native oracle build failed (%s); portable x86-64-v3 build" % e, file=sys.stderr, flush=True)
        march = "x86-64-v3"
    orc = cases.oracle_case(p)
    st = orc.new_state()
    for k, q in zip("uvwp", init):
        st[k][1:-1, 1:-1, 1:-1] = q
    orc.init_bounds(st)
    orc.rk3_step(st, DT)  # warm-up (page-in, FFT tables)
    times = []
    t_all = time.perf_counter()
    while True:
        t0 = time.perf_counter()
        orc.rk3_step(st, DT)
        times.append(time.perf_counter() - t0)
        print("[cpu] oracle step %d: %.2f s" % (len(times), times[-1]), file=sys.stderr, flush=True)
        if time.perf_counter() - t_all > budget_s or len(times) >= 10:
            break
    n = int(np.prod(p["ng"]))
    med = float(np.median(times))
    model = "unknown"
    try:
        with open("/proc/cpuinfo") as f:
            model = next(ln.split(":", 1)[1].strip() for ln in f if ln.startswith("model name"))
    except (OSError, StopIteration):
        pass
    return {"value": med / n * 1e9, "unit": "ns/cell/step", "cores": O.lib().orc_num_threads(),
            "kind": "port",
            "sample": "%d full RK3 steps of the %dx%dx%d channel (median), oracle C/OpenMP, FP64"
                      % (len(times), *p["ng"]),
            "cpu_model": model, "nproc": len(os.sched_getaffinity(0)), "os_cpu_count": os.cpu_count(),
            "omp_num_threads": os.environ.get("OMP_NUM_THREADS"), "march": march,
            "min_s": float(np.min(times)), "max_s": float(np.max(times)), "avg_s": float(np.mean(times))}


def run(p, init, precision, args, sweeps=0, icheck=1, solves=True):
    """Time K steps (and, with solves, S Poisson solves) in one precision mode; return figures and
    final fields.  sweeps: FP64 refinement sweeps of the last substep's pressure solve (mixed mode),
    on every icheck-th step: the steps whose divergence the reference checks (main.f90:501-510,
    icheck = 10 in every example input), counted so that the last timed step is one of them."""
    import numpy as np
    import torch
    from cans_amd import _lib as LB
    from cans_amd.host import Simulation
    from cans_amd import initflow as IF
    from tests import cases
    label = ("mixed, %d refinement sweeps every %d steps" % (sweeps, icheck)) if precision == LB.CANS_MIXED \
        else "fp64"
    print("[bench] %s: set-up" % label, file=sys.stderr, flush=True)
    sim = Simulation(cases.host_case(p), precision=precision)

    def refine(i):  # timed step i: refine the last substep, whose projection sets div(u), on icheck steps
        if sweeps:
            on = (args.steps - 1 - i) % icheck == 0
            sim.set_refinement(sweeps if on else 0, 4 if on else 0)

    sim.set_state(*(IF.halo(p["ng"], q) for q in init))
    sim.init_bounds()
    ncell = int(np.prod(p["ng"]))
    for i in range(args.warmup):
        refine(i - args.warmup)
        sim.rk3_step(DT)
    torch.cuda.synchronize()
    # the context runs on torch's current stream, so these events bracket its work
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    t0 = time.perf_counter()
    e0.record()
    for i in range(args.steps):
        refine(i)
        sim.rk3_step(DT)
    e1.record()
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    ms_step = e0.elapsed_time(e1) / args.steps
    print("[bench] %s: %.3f ms/step" % (label, ms_step), file=sys.stderr, flush=True)
    final = {k: getattr(sim, k).interior() for k in "uvwp"}
    tot, divmax = sim.chkdiv()
    res = {"ms_per_step": ms_step, "ns_per_cell_step": ms_step * 1e-3 / ncell * 1e9, "wall_s": wall,
           "divmax": divmax}
    if precision == LB.CANS_MIXED:
        res["refinement_sweeps"] = sweeps
        if sweeps:
            res["refined_every_steps"] = icheck
    if not solves:
        sim.close()
        del sim
        torch.cuda.empty_cache()
        return res, final
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
    b_p = 5 * 2 * sim.ctx.p_es * ncell
    gbs = b_p / t_p / 1e9
    refined = None
    if precision == LB.CANS_MIXED:  # FP64 solves from FP32 inner solves and two refinement sweeps
        f = sim.ctx.field()
        ev = [torch.cuda.Event(enable_timing=True) for _ in range(4)]

        def loop(e0, e1, solve):
            e0.record()
            for _ in range(args.solves):
                f.buf.copy_(sim.u.buf)  # any FP64 right-hand side takes the same time
                if solve:
                    sim.solver_refined(f, 2)
            e1.record()

        loop(ev[0], ev[1], True)
        torch.cuda.synchronize()
        loop(ev[0], ev[1], True)
        loop(ev[2], ev[3], False)
        torch.cuda.synchronize()
        ms = (ev[0].elapsed_time(ev[1]) - ev[2].elapsed_time(ev[3])) / args.solves
        refined = {"sweeps": 2, "ms": ms,
                   "note": "FP64 Poisson solve: three FP32-storage inner solves, two FP64 residuals and sums"}
        del f
    sim.close()
    del sim
    torch.cuda.empty_cache()
    res["poisson"] = {"ms": t_p * 1e3, "GBps": gbs, "pct_hbm_peak": 100.0 * gbs / HBM_PEAK_GBS,
                      "algorithmic_bytes": b_p}
    if refined:
        res["poisson_refined"] = refined
    return res, final


def solve_only(p, precision, args):
    """args.solves Poisson solves of the channel's pressure system on a seeded rhs, timed with
    HIP events on the context stream (the same figure as the step runs' "poisson")."""
    import numpy as np
    import torch
    from cans_amd.host import Simulation
    from tests import cases
    sim = Simulation(cases.host_case(p), precision=precision)
    sim.pp.buf.uniform_(-1, 1)
    sim.solver()
    torch.cuda.synchronize()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record()
    for _ in range(args.solves):
        sim.solver()
    e1.record()
    torch.cuda.synchronize()
    t_p = e0.elapsed_time(e1) / args.solves * 1e-3
    b_p = 5 * 2 * sim.ctx.p_es * int(np.prod(p["ng"]))
    sim.close()
    del sim
    torch.cuda.empty_cache()
    gbs = b_p / t_p / 1e9
    print("[bench] solve-only: %.3f ms (%.0f GB/s)" % (t_p * 1e3, gbs), file=sys.stderr, flush=True)
    return {"ms": t_p * 1e3, "GBps": gbs, "pct_hbm_peak": 100.0 * gbs / HBM_PEAK_GBS, "algorithmic_bytes": b_p}


def library_fft_reference(p, reps=10):
    """The vendor library on the solve's x and y transforms, for comparison with the four transform passes
    of the Poisson solve: rocFFT through torch.fft (hipFFT) on the channel's (n3, n2, n1) field, a real
    transform along x (rfft, n1/2+1 complex), a complex one along y, then their inverses; FP32 and FP64,
    out of place, mean of `reps` after a warm-up."""
    import torch
    n1, n2, n3 = p["ng"]
    out = {"workload": "rfft along x + fft along y, then ifft along y + irfft along x, on (n3, n2, n1) = "
                       "(%d, %d, %d), torch.fft (rocFFT)" % (n3, n2, n1)}
    for name, dt in (("fp32_ms", torch.float32), ("fp64_ms", torch.float64)):
        x = torch.empty((n3, n2, n1), dtype=dt, device="cuda").uniform_(-1, 1)

        def run():
            c = torch.fft.fft(torch.fft.rfft(x, dim=2), dim=1)
            return torch.fft.irfft(torch.fft.ifft(c, dim=1), n=n1, dim=2)

        y = run()
        torch.cuda.synchronize()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record()
        for _ in range(reps):
            y = run()
        e1.record()
        torch.cuda.synchronize()
        out[name] = e0.elapsed_time(e1) / reps
        out[name.replace("_ms", "_roundtrip_rel_err")] = float((y - x).abs().max() / x.abs().max())
        del x, y
        torch.cuda.empty_cache()
    print("[bench] rocFFT x/y transforms: %.3f ms FP32, %.3f ms FP64" % (out["fp32_ms"], out["fp64_ms"]),
          file=sys.stderr, flush=True)
    return out


def secondary_cavity(args, precision="fp64"):
    """A second workload beside the headline: the 256^3 heated cube (cases.params("dhc3d")) with no-slip
    walls on all faces, a scalar with buoyancy and 3-D implicit diffusion (is_impdiff: per substep the
    scalar's Helmholtz system, then the three velocity systems, whose transforms along each component's
    wall-normal direction are face-centred DST-I), a DCT x DCT pressure solve, fixed dt.  FP64, or the
    mixed mode (FP32 pressure side; the velocity and its Helmholtz solves FP64)."""
    import numpy as np
    import torch
    from cans_amd import _lib as LB
    from cans_amd import initflow as IF
    from cans_amd.host import Simulation
    from tests import cases
    p = cases.params("dhc3d")
    p["is_impdiff_1d"] = False
    p["is_impdiff"] = True
    print("[bench] heated cube %s, %s: set-up" % (p["ng"], precision), file=sys.stderr, flush=True)
    sim = Simulation(cases.host_case(p), precision=LB.CANS_MIXED if precision == "mixed" else LB.CANS_FP64)
    u, v, w, pr = cases.initial_state("dhc3d", p, sim.grid, seed=3)
    sc = cases.initial_scalars("dhc3d", p, sim.grid)
    sim.set_state(*(IF.halo(p["ng"], q) for q in (u, v, w, pr)), s=[IF.halo(p["ng"], q) for q in sc])
    sim.init_bounds()
    dt = 1e-3
    for _ in range(2):
        sim.rk3_step(dt)
    torch.cuda.synchronize()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    nst = max(1, args.steps // 2)
    e0.record()
    for _ in range(nst):
        sim.rk3_step(dt)
    e1.record()
    torch.cuda.synchronize()
    ms = e0.elapsed_time(e1) / nst
    _, divmax = sim.chkdiv()
    ncell = int(np.prod(p["ng"]))
    # the three velocity Helmholtz systems of one substep (alpha = -visc*dtrk/2), their x and y
    # transforms differing along each component's wall normal: as one batch (one launch per pass
    # for all three) and one cans_solve_helmholtz call at a time
    fields = [sim.ctx.field() for _ in range(3)]
    for f, q in zip(fields, (sim.u, sim.v, sim.w)):
        f.buf.copy_(q.buf)
    alpha = -0.5 * sim.grid.visc * (32. / 60.) * dt
    helm = {}
    for form, fn in (("batch", lambda: sim.solve_helmholtz_batch(fields, sim.hplans, [alpha] * 3)),
                     ("one_at_a_time", lambda: [hp.solve_helmholtz(f, alpha) for f, hp in zip(fields, sim.hplans)])):
        fn()
        torch.cuda.synchronize()
        e0.record()
        for _ in range(nst):
            fn()
        e1.record()
        torch.cuda.synchronize()
        helm["ms_" + form] = e0.elapsed_time(e1) / nst
    del fields
    sim.close()
    del sim
    torch.cuda.empty_cache()
    print("[bench] heated cube %s: %.3f ms/step; velocity Helmholtz batch %.3f ms, one at a time %.3f ms"
          % (precision, ms, helm["ms_batch"], helm["ms_one_at_a_time"]), file=sys.stderr, flush=True)
    return {"workload": "heated cube %dx%dx%d, no-slip walls, scalar with buoyancy, 3-D implicit diffusion "
                        "(is_impdiff; velocity Helmholtz with face-centred DST-I along the wall normal), "
                        "DCT x DCT pressure, dt=1e-3, %s" % (tuple(p["ng"]) + (precision,)),
            "ms_per_step": ms, "ns_per_cell_step": ms * 1e-3 / ncell * 1e9, "steps": nst, "divmax": divmax,
            "velocity_helmholtz": helm}


def secondary_boxes(args):
    """64 independent Poisson boxes (nx, ny from {32, 48, 64, 96, 128}, nz uniform in [32, 160],
    tanh stretching in [0, 2], periodic x and y, Neumann z, seed 2024) solved as one batch
    (cans_poisson_batch_solve: five launches for all boxes), FP64 and mixed storage, beside
    the same boxes solved one cans_solver call at a time (five launches per box)."""
    import numpy as np
    import torch
    from cans_amd import _lib as LB
    from cans_amd import initflow as IF
    from cans_amd.host import Context, Grid, Plan, PoissonBatch
    from tests import cases
    boxes = cases.poisson_boxes(64, seed=2024)
    ncell = sum(int(np.prod(b["n"])) for b in boxes)
    big = tuple(max(b["n"][d] for b in boxes) for d in range(3))
    nrep = max(5, args.solves)

    def timed(fn, st=None):  # events on st (the stream the work runs on), or torch's current one
        fn()
        torch.cuda.synchronize()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record(st)
        for _ in range(nrep):
            fn()
        e1.record(st)
        torch.cuda.synchronize()
        return e0.elapsed_time(e1) / nrep

    out = {"workload": "64 independent Poisson boxes, nx,ny in {32,48,64,96,128}, nz in [32,160], "
                       "stretching gr in [0,2], PP x/y, NN z (seed 2024)", "cells": ncell, "boxes": len(boxes)}
    for name, prec, es in (("fp64", LB.CANS_FP64, 8), ("mixed", LB.CANS_MIXED, 4)):
        print("[bench] %d Poisson boxes (%d cells), %s: set-up" % (len(boxes), ncell, name), file=sys.stderr,
              flush=True)
        stream = torch.cuda.Stream()  # the context's own stream, so that a solve can be captured
        ctx = Context(big, precision=prec, stream=stream)
        batch = PoissonBatch(ctx, boxes)
        rng = np.random.default_rng(6)
        for i, b in enumerate(boxes):
            batch.set(i, IF.halo(b["n"], cases.box_rhs(b, 1 / batch.dzfi[i][1:-1], rng)))
        torch.cuda.synchronize()  # the copies ran on torch's stream, the solves run on the context's
        ms = timed(batch.solve, stream)
        nbytes = 10 * es * ncell  # 5 passes x (s_in + s_out) per cell
        out[name] = {"ms_per_batch_solve": ms, "algorithmic_bytes": nbytes, "GBps": nbytes / (ms * 1e-3) / 1e9,
                     "frac_hbm_peak": nbytes / (ms * 1e-3) / 1e9 / HBM_PEAK_GBS, "threads": batch.threads()}
        # the same solve captured once in a HIP graph (its five launches as one replay)
        g = torch.cuda.CUDAGraph()
        torch.cuda.synchronize()
        with torch.cuda.graph(g, stream=stream):
            batch.solve()
        # replay() launches the graph on torch's current stream: time it there
        out[name]["ms_per_batch_solve_graph_replay"] = timed(g.replay)
        del g
        batch.close()
        del batch
        if name == "fp64":  # the same boxes, one cans_solver call each
            single = []
            for b in boxes:
                p = dict(ng=b["n"], l=b["l"], gtype=1, gr=b["gr"], cbcvel="PP" * 9, cbcpre=b["cbc"], visci=100.0)
                c1 = Context(b["n"], precision=prec)
                pl = Plan(c1, Grid(cases.host_case(p)), b["cbc"], np.zeros(6))
                f = c1.field(kind=LB.CANS_FIELD_PRESSURE)
                single.append((c1, pl, f))
            ms1 = timed(lambda: [pl.solve(f) for _, pl, f in single])
            out[name]["ms_one_box_at_a_time"] = ms1
            for c1, pl, f in single:
                pl.close()
                c1.close()
            del single
        ctx.close()
        torch.cuda.empty_cache()
        print("[bench] boxes %s: %.3f ms per batch solve (%.0f GB/s)" % (name, ms, out[name]["GBps"]),
              file=sys.stderr, flush=True)
    return out


ADJOINT_WORKLOADS = {
    "tgv": dict(ng=(256, 256, 256), seed=1, dt=5e-4, nsteps=32, every=4,
                text="gradient of the kinetic energy after 32 RK3 steps, Taylor-Green 256^3, nu=1/1600, "
                     "dt=5e-4, FP64, checkpoints every 4 steps"),
    "channel": dict(ng=(512, 256, 192), seed=3, dt=2e-4, nsteps=16, every=4,
                    text="gradient of the kinetic energy after 16 RK3 steps, channel 512x256x192 (walls in z, "
                         "tanh 1.8, constant flow rate), Re_b=5600, dt=2e-4, FP64, checkpoints every 4 steps"),
}


def secondary_adjoint(args, name="tgv"):
    """Gradient of the kinetic energy after nsteps RK3 steps w.r.t. the initial velocity (FP64): forward with
    checkpoints every 4 steps, backward with recompute and the adjoint substeps (cans_amd.adjoint).  tgv: the
    256^3 Taylor-Green flow (nu = 1/1600, dt = 5e-4, 32 steps); channel: 512x256x192 with walls in z on the
    tanh-stretched grid and bulk forcing (dt = 2e-4, 16 steps)."""
    import numpy as np
    import torch
    from cans_amd.adjoint import kinetic_energy_after_steps
    from cans_amd.host import Simulation
    from tests import cases
    wl = ADJOINT_WORKLOADS[name]
    p = cases.params(name, ng=wl["ng"])
    print("[bench] adjoint %s %s: set-up" % (name, p["ng"]), file=sys.stderr, flush=True)
    sim = Simulation(cases.host_case(p))
    u0 = list(cases.initial_state(name, p, sim.grid, seed=wl["seed"])[:3])
    if p["cbcpre"][4:6] != "PP":
        u0[2][:, :, -1] = 0.  # w's wall face
    dt, nsteps, every = wl["dt"], wl["nsteps"], wl["every"]
    res = {}
    for rep in range(2):  # the first run warms up (allocations, code loading)
        t0 = [torch.tensor(np.ascontiguousarray(a.transpose(2, 1, 0)), device="cuda", requires_grad=True)
              for a in u0]
        torch.cuda.synchronize()
        e = [torch.cuda.Event(enable_timing=True) for _ in range(3)]
        e[0].record()
        E = kinetic_energy_after_steps(sim, *t0, dt, nsteps, every)
        e[1].record()
        E.backward()
        e[2].record()
        torch.cuda.synchronize()
        res = {"forward_ms": e[0].elapsed_time(e[1]), "backward_ms": e[1].elapsed_time(e[2])}
    ncell = int(np.prod(p["ng"]))
    sim.close()
    del sim, t0, E
    torch.cuda.empty_cache()
    out = {"workload": wl["text"], "steps": nsteps,
           "forward_ms": res["forward_ms"], "backward_ms": res["backward_ms"],
           "backward_over_forward": res["backward_ms"] / res["forward_ms"],
           "ns_per_cell_step_forward": res["forward_ms"] * 1e-3 / nsteps / ncell * 1e9}
    print("[bench] adjoint %s: forward %.1f ms, backward %.1f ms" % (name, res["forward_ms"], res["backward_ms"]),
          file=sys.stderr, flush=True)
    return out


def copy_bandwidth():
    """Device copy of a known size: the box's streaming bandwidth."""
    import torch
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
    gbs = 2 * nbytes * 5 / (c0.elapsed_time(c1) * 1e-3) / 1e9
    del src, dst
    torch.cuda.empty_cache()
    return gbs


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--ng", type=int, nargs=3, default=(1024, 512, 384))
    ap.add_argument("--precision", choices=("both", "mixed", "fp64"), default="both")
    ap.add_argument("--solves", type=int, default=10)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-secondary", action="store_true")
    ap.add_argument("--boxes-only", action="store_true", help="time only the 64 Poisson boxes (for profiling)")
    ap.add_argument("--adjoint-only", nargs="?", const="both", choices=("both", "tgv", "channel"),
                    help="time only the adjoint workloads, both or one of them (for profiling)")
    ap.add_argument("--cavity-only", action="store_true", help="time only the 256^3 heated cube (for profiling)")
    ap.add_argument("--solve-only", action="store_true",
                    help="time only the channel's Poisson solves, FP32 storage then FP64 (for profiling: after each "
                         "mode's set-up, the last dispatches of every pass kernel are the timed solves)")
    ap.add_argument("--dtdma-slabs", type=int, default=0,
                    help="with --solve-only: the z elimination as the distributed solve over this many z slabs "
                         "(is_poisson_dtdma; 0: the plain sweep)")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py: this build runs on one GPU; multi-GPU runs (--gpus %d) are not part of it" % args.gpus,
              file=sys.stderr)
        sys.exit(2)

    only = [f for flag, f in ((args.cavity_only, secondary_cavity), (args.boxes_only, secondary_boxes),
                              (args.adjoint_only in ("both", "tgv"), secondary_adjoint),
                              (args.adjoint_only in ("both", "channel"), lambda a: secondary_adjoint(a, "channel")))
            if flag]
    if only:
        print(json.dumps({"secondary": [f(args) for f in only]}))
        return

    import numpy as np
    from cans_amd import _lib as LB
    from cans_amd.host import Grid
    from tests import cases

    p = cases.params("channel", ng=tuple(args.ng))
    if args.solve_only:
        p["dtdma_slabs"] = args.dtdma_slabs
        print(json.dumps({"poisson": {m: solve_only(p, prec, args) for m, prec in
                                      (("mixed", LB.CANS_MIXED), ("fp64", LB.CANS_FP64))}}))
        return
    ncell = int(np.prod(p["ng"]))
    grid = Grid(cases.host_case(p))
    # one seeded state, FP32-representable so that both modes start from the same values
    init = [q.astype(np.float32).astype(np.float64) for q in cases.initial_state("channel", p, grid, seed=1)]
    # mixed: config 2 as stated, two FP64 refinement sweeps of the last substep's pressure solve on
    # the steps whose divergence is checked (every 10th, the last timed step among them);
    # mixed_refined_every_step: the sweeps on every step; mixed_plain: a single FP32-storage solve
    modes = {"both": ("mixed", "mixed_refined_every_step", "mixed_plain", "fp64"),
             "mixed": ("mixed", "mixed_refined_every_step", "mixed_plain"), "fp64": ("fp64",)}[args.precision]
    res, final = {}, {}
    for m in modes:
        res[m], final[m] = run(p, init, LB.CANS_FP64 if m == "fp64" else LB.CANS_MIXED, args,
                               sweeps=0 if m in ("fp64", "mixed_plain") else 2,
                               icheck=ICHECK if m == "mixed" else 1, solves=m == "mixed" or m == "fp64")
    head = res[modes[0]]
    head_mode = modes[0]
    wl = ("periodic channel %dx%dx%d, stretched z (gr=1.8), Re_b=5600, dt=2e-4; %s"
          % (tuple(p["ng"]) + ("mixed precision (FP32 pressure-side storage and FFTs, FP64 velocity, stencils, "
                               "z elimination and divergence, 2 FP64 refinement sweeps of the last substep's "
                               "pressure solve on the divergence-checked steps, every %d) vs FP64" % ICHECK
                               if head_mode == "mixed" else head_mode,)))
    pois = head["poisson"]
    out = {
        "metric": METRIC,
        "value": head["ns_per_cell_step"],
        "unit": "ns/cell/step",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": head["ms_per_step"],
        "higher_is_better": False,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64/f32" if head_mode == "mixed" else "f64",
        "data": "synthetic (seeded laminar profile + 10% noise)",
        "config": {"workload": wl, "cells": ncell, "precision": head_mode,
                   "storage": ("velocity f64; p, pp, RHS_old, RK increment and transform passes f32"
                               if head_mode == "mixed" else "f64")},
        "poisson": pois,
        "roofline": {"bound": "hbm", "achieved": pois["GBps"], "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": pois["GBps"] / HBM_PEAK_GBS,
                     "traffic": SOLVE_TRAFFIC_BYTES.get((tuple(p["ng"]), head_mode)),
                     "traffic_source": TRAFFIC_SOURCE if (tuple(p["ng"]), head_mode) in SOLVE_TRAFFIC_BYTES else None,
                     "algorithmic_bytes": pois["algorithmic_bytes"],
                     "kernel": "Poisson solve (x-fwd, y-fwd, z-Thomas, y-bwd, x-bwd passes), %s storage"
                               % ("f32" if head_mode == "mixed" else "f64")},
    }
    for m in modes:
        out[m] = res[m]
    if "fp64" in modes and "mixed" in modes:
        out["speedup_mixed_vs_fp64"] = res["fp64"]["ms_per_step"] / res["mixed"]["ms_per_step"]
        out["speedup_mixed_plain_vs_fp64"] = res["fp64"]["ms_per_step"] / res["mixed_plain"]["ms_per_step"]
        out["speedup_mixed_refined_every_step_vs_fp64"] = (res["fp64"]["ms_per_step"] /
                                                           res["mixed_refined_every_step"]["ms_per_step"])
        b = final["fp64"]
        wgt = np.broadcast_to(1 / grid.dzfi[1:-1], b["p"].shape)
        db = b["p"] - (b["p"] * wgt).sum() / wgt.sum()
        for m in ("mixed", "mixed_refined_every_step", "mixed_plain"):
            err = {}
            a = final[m]
            for k in "uvw":
                err[k] = float(np.abs(a[k] - b[k]).max() / np.abs(b[k]).max())
            da = a["p"] - (a["p"] * wgt).sum() / wgt.sum()
            err["p_gauge_fixed"] = float(np.abs(da - db).max() / np.abs(db).max())
            err["steps"] = args.warmup + args.steps
            out[m + "_vs_fp64_rel_linf"] = err
    del final
    out["copy_GBps"] = copy_bandwidth()
    out["library_fft_reference"] = library_fft_reference(p)
    if not args.no_secondary:
        out["secondary"] = [secondary_cavity(args), secondary_cavity(args, "mixed"), secondary_boxes(args),
                            secondary_adjoint(args), secondary_adjoint(args, "channel")]
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline(p, init)
    print(json.dumps(out))


if __name__ == "__main__":
    main()
