"""The reference's lid-driven-cavity regression (tests/lid_driven_cavity) replayed on the GPU.

2x64x64 cells, Re = 1000, lid v = 1, adaptive dt from chkdt every 10 steps, 1500 fused RK3 steps;
v(1, 33, 1:64) must match the reference's committed fixture at its own tolerance (rtol 1e-7).
"""
import os

import numpy as np
import pytest

from tests import cases

pytestmark = pytest.mark.gpu
HERE = os.path.dirname(os.path.abspath(__file__))


def test_ldc_fixture_on_gpu():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    from cans_amd.host import Simulation
    print("\n[case] ldc 2x64x64 1500 steps", flush=True)
    p = cases.params("ldc")
    sim = Simulation(cases.host_case(p))
    sim.init_bounds()
    cfl, dtmax = 0.95, 1e5
    dt = min(cfl * sim.chkdt(), dtmax)
    for istep in range(1, 1501):
        sim.rk3_step(dt)
        if istep % 10 == 0:
            dt = min(cfl * sim.chkdt(), dtmax)
            divtot, divmax = sim.chkdiv()
            assert divmax < 2.2e-9, (istep, divmax)
    ref = np.loadtxt(os.path.join(HERE, "golden", "data_ldc_re1000.txt"))
    got = sim.v.numpy()[1, 33, 1:65]
    err = np.abs(got / ref[:, 1] - 1).max()
    print("ldc max rel err vs fixture %.3e (dt %.6f)" % (err, dt), flush=True)
    np.testing.assert_allclose(got, ref[:, 1], rtol=1e-7, atol=0)
