"""The mixed-precision channel at its stated size, 1024 x 512 x 384, against the FP64 oracle.

Stretched z (gr = 1.8), periodic x (4 pi) and y (2 pi), no-slip walls, constant flow
rate, Re_b = 5600, dt = 2e-4; a seeded laminar profile with 10 % noise (FP32-representable,
so that both sides start from the same values).  The GPU runs the mixed mode (FP32
transforms and pressure-side storage, FP64 velocity, stencils, z elimination and
divergence) with two FP64 refinement sweeps of the last substep's pressure solve; the
oracle (C/OpenMP restatement of the reference's FP64 path) runs the same 10 RK3 steps.
Bar: relative L-inf error <= 1e-6 for u, v, w and the gauge-fixed pressure, and the max
divergence below the reference's FP64 abort threshold `small` (param.f90:18).
"""
import time

import numpy as np
import pytest

from cans_amd import initflow as IF
from tests import cases

pytestmark = pytest.mark.gpu

NG = (1024, 512, 384)
STEPS = 10
DT = 2e-4
SMALL = 2.220446049250313e-9


def test_mixed_channel_1024x512x384_vs_fp64_oracle():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    from cans_amd import _lib as LB
    from cans_amd.host import Simulation
    p = cases.params("channel", ng=NG)
    print("\n[case] mixed channel %s, %d steps, 2 refinement sweeps on substep 3" % (NG, STEPS), flush=True)
    sim = Simulation(cases.host_case(p), precision=LB.CANS_MIXED)
    sim.set_refinement(2, 4)
    init = [q.astype(np.float32).astype(np.float64) for q in cases.initial_state("channel", p, sim.grid, seed=1)]
    sim.set_state(*(IF.halo(NG, q) for q in init))
    sim.init_bounds()
    div = []
    for it in range(STEPS):
        sim.rk3_step(DT)
        div.append(sim.chkdiv()[1])
    sim.ctx.sync()
    print("gpu: %d steps done, divmax per step %s" % (STEPS, ["%.2e" % d for d in div]), flush=True)
    gpu = {k: getattr(sim, k).interior() for k in "uvwp"}
    dzf = 1 / sim.grid.dzfi[1:-1]
    sim.close()
    del sim
    torch.cuda.empty_cache()
    orc = cases.oracle_case(p)
    st = orc.new_state()
    for k, q in zip("uvwp", init):
        st[k][1:-1, 1:-1, 1:-1] = q
    del init
    orc.init_bounds(st)
    for it in range(STEPS):
        t0 = time.perf_counter()
        orc.rk3_step(st, DT)
        print("oracle step %d: %.1f s" % (it + 1, time.perf_counter() - t0), flush=True)
    err = {}
    for k in "uvw":
        b = st[k][1:-1, 1:-1, 1:-1]
        err[k] = float(np.abs(gpu[k] - b).max() / np.abs(b).max())
    wgt = np.broadcast_to(dzf, gpu["p"].shape)
    a, b = gpu["p"], st["p"][1:-1, 1:-1, 1:-1]
    da = a - (a * wgt).sum() / wgt.sum()
    db = b - (b * wgt).sum() / wgt.sum()
    err["p"] = float(np.abs(da - db).max() / np.abs(db).max())
    _, omx = orc.chkdiv(st["u"], st["v"], st["w"])
    print("rel L-inf vs FP64 oracle: %s; divmax gpu %.3e oracle %.3e"
          % (", ".join("%s %.3e" % kv for kv in err.items()), div[-1], omx), flush=True)
    for k, e in err.items():
        assert e <= 1e-6, (k, e)
    assert max(div) < SMALL
