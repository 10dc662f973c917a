"""Radix-4 power-of-two FFT stages (cans_fft_set_radix; the launch-shape tuner may choose them
over radix 8): the same transforms with another split into stages, so solves through them match
the oracle at the FP64 bar (1e-12) and the mixed mode's 1e-6, and the radix-8 solve to rounding.
Lengths cover an odd number of radix-2 factors (a final radix-2 stage), radix-3/5 stages after
the power-of-two ones, and every x/y kind."""
import ctypes as C

import numpy as np
import pytest

from cans_amd import initflow as IF
from tests import cases

pytestmark = pytest.mark.gpu

CASES = [("channel", (64, 32, 48)), ("channel", (128, 96, 16)), ("cavity_dd", (32, 40, 12)),
         ("open_nd", (16, 24, 12)), ("box", (24, 32, 10)), ("open_x", (64, 16, 8))]


@pytest.fixture(scope="module")
def torch():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    return torch


def gauge(a, dzf):
    wgt = np.broadcast_to(dzf, a.shape)
    return a - (a * wgt).sum() / wgt.sum()


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
@pytest.mark.parametrize("name,ng", CASES, ids=["%s-%dx%dx%d" % ((c,) + g) for c, g in CASES])
def test_radix4_stages_match_oracle(torch, name, ng, precision):
    from cans_amd import _lib as LB
    from cans_amd.host import Simulation
    print("\n[case] radix 4 %s %s %s" % (name, ng, precision), flush=True)
    p = cases.params(name, ng=ng)
    prec = LB.CANS_MIXED if precision == "mixed" else LB.CANS_FP64
    sims = [Simulation(cases.host_case(p), precision=prec) for _ in range(2)]
    for sim, r in zip(sims, (8, 4)):
        assert LB.lib().cans_fft_set_radix(sim.plan.h, (C.c_int * 4)(r, r, r, r)) == LB.CANS_OK
        got = (C.c_int * 4)()
        LB.check(LB.lib().cans_fft_radix(sim.plan.h, got), "cans_fft_radix")
        assert list(got) == [r] * 4
    rng = np.random.default_rng(4)
    rhs = rng.uniform(-1, 1, tuple(ng))
    dzf = 1 / sims[0].grid.dzfi[1:-1]
    singular = "D" not in p["cbcpre"]
    if singular:
        rhs = gauge(rhs, dzf)
    if precision == "mixed":
        rhs = rhs.astype(np.float32).astype(np.float64)
    out = []
    for sim in sims:
        sim.pp.set(IF.halo(ng, rhs))
        sim.solver()
        sim.ctx.sync()
        out.append(sim.pp.interior())
    orc = cases.oracle_case(p)
    st = orc.new_state()
    st["pp"][...] = IF.halo(ng, rhs)
    orc.solver(st["pp"])
    ref = st["pp"][1:-1, 1:-1, 1:-1]
    if singular:
        out = [gauge(o, dzf) for o in out]
        ref = gauge(ref, dzf)
    scale = np.abs(ref).max()
    err4 = np.abs(out[1] - ref).max() / scale
    err84 = np.abs(out[1] - out[0]).max() / scale
    print("radix-4 solve vs oracle %.3e, vs radix 8 %.3e" % (err4, err84), flush=True)
    tol = 1e-6 if precision == "mixed" else 1e-12
    assert err4 < tol and err84 < tol


def test_radix_setting_checked(torch):
    from cans_amd import _lib as LB
    from cans_amd.host import Simulation
    sim = Simulation(cases.host_case(cases.params("channel", ng=(32, 16, 8))))
    assert LB.lib().cans_fft_set_radix(sim.plan.h, (C.c_int * 4)(8, 16, 8, 8)) == LB.CANS_EINVAL
