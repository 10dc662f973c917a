"""A fused RK3 step captured in a HIP graph and replayed: every replay must give the
bit-identical fields of the same step run eagerly (same kernels, same order), also with
scalars, buoyancy and implicit diffusion.  Eager work
and synchronisations run between the replays, and each replay is checked, so a capture
that records a memset or a stale argument shows up on a later replay."""
import numpy as np
import pytest

from cans_amd import initflow as IF
from tests import cases

pytestmark = pytest.mark.gpu


GRAPH_CASES = [("channel", "fp64"), ("channel", "mixed"), ("rb_channel", "fp64-impdiff")]


@pytest.mark.parametrize("name,precision", GRAPH_CASES, ids=["%s-%s" % c for c in GRAPH_CASES])
def test_graph_replays_match_eager(name, precision):
    """channel: the velocity-only step; rb_channel: two scalars (one forced), buoyancy and
    implicit z diffusion (batched Helmholtz solves) in the captured step."""
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    from cans_amd import _lib as LB
    from cans_amd.host import Simulation
    print("\n[case] graph replay %s %s" % (name, precision), flush=True)
    prec = LB.CANS_MIXED if precision == "mixed" else LB.CANS_FP64
    ng = (64, 32, 48) if name == "channel" else (32, 16, 24)
    p = cases.params(name, ng=ng)
    if precision.endswith("impdiff"):
        p["is_impdiff_1d"] = True
    grid = Simulation(cases.host_case(p), precision=prec).grid
    init = cases.initial_state(name, p, grid, seed=3)
    sinit = cases.initial_scalars(name, p, grid, seed=4)
    dt = 2e-4
    fields = "uvwp"

    def fresh(stream):
        sim = Simulation(cases.host_case(p), precision=prec, stream=stream)
        sim.set_state(*(IF.halo(p["ng"], q) for q in init), s=[IF.halo(p["ng"], q) for q in sinit])
        torch.cuda.synchronize()  # the uploads ran on torch's current stream
        sim.init_bounds()
        return sim

    def snapshot(sim):
        out = {k: getattr(sim, k).numpy() for k in fields}
        for i, f in enumerate(sim.s):
            out["s%d" % i] = f.numpy()
        return out

    # eager reference: 1 warm-up step + 3 steps, fields after each
    eager = fresh(None)
    eager.rk3_step(dt)
    ref = []
    for _ in range(3):
        eager.rk3_step(dt)
        eager.ctx.sync()
        ref.append(snapshot(eager))
    # graph: the same warm-up step eagerly, then one captured step replayed 3 times
    s = torch.cuda.Stream()
    sim = fresh(s)
    with torch.cuda.stream(s):
        sim.rk3_step(dt)
    sim.ctx.sync()
    g = sim.capture_step(dt)
    for r in range(3):
        g.replay()
        torch.cuda.synchronize()
        junk = torch.ones(1 << 20, device="cuda")  # eager work between replays
        junk.mul_(2.0).sum().item()
        got_all = snapshot(sim)
        for k, got in got_all.items():
            assert np.array_equal(got, ref[r][k]), (name, precision, r, k, float(np.abs(got - ref[r][k]).max()))
        print("replay %d identical to eager" % r, flush=True)


def test_plan_tuner_shapes():
    """cans_fftini times the pass launch shapes and keeps valid ones for the storage types in use."""
    import ctypes as C
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    from cans_amd import _lib as LB
    from cans_amd.host import Simulation
    for prec in (LB.CANS_FP64, LB.CANS_MIXED):
        p = cases.params("channel", ng=(64, 32, 48))
        sim = Simulation(cases.host_case(p), precision=prec)
        sh = (C.c_int * 4)()
        LB.check(LB.lib().cans_fft_tuning(sim.plan.h, sh), "cans_fft_tuning")
        xf, xd, yf, yd = list(sh)
        print("[tuner] precision %d: x %d/%d y %d/%d" % (prec, xf, xd, yf, yd), flush=True)
        assert xd in (128, 256, 512) and yd in (4, 8, 16, 32)
        if prec == LB.CANS_MIXED:
            assert xf in (128, 256, 512) and yf in (4, 8, 16, 32)
        else:
            assert xf == 0 and yf == 0
        sim.close()
