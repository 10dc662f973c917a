"""Replays the reference's own regression test tests/lid_driven_cavity with the oracle.

Case parameters restate tests/lid_driven_cavity/input.nml: 2x64x64 cells,
l = (1/32, 1, 1), uniform grid, Re = 1000, lid v = 1 on the upper z wall,
adaptive dt (cfl 0.95, recomputed every icheck = 10 steps), 1500 steps.
The fixture tests/golden/data_ldc_re1000.txt is the reference's committed
data file; the reference compares v(1, ny/2+1, :) against it at rtol 1e-7
(tests/lid_driven_cavity/test.py:5-8).
"""
import os

import numpy as np

from oracle import oracle as O

HERE = os.path.dirname(os.path.abspath(__file__))


def ldc_case():
    bcvel = np.zeros(18)
    bcvel[1 + 2 * 2 + 6 * 1] = 1.0  # bcvel(1,3,2) = 1
    return O.OracleCase(ng=(2, 64, 64), l=(0.03125, 1.0, 1.0), gtype=1, gr=0.0,
                        cbcvel="PPDDDD" * 3, cbcpre="PPNNNN", bcvel=bcvel, visci=1000.0)


def test_ldc_fixture_replay():
    case = ldc_case()
    st = case.new_state()
    case.init_bounds(st)
    cfl, dtmax = 0.95, 1e5
    dt = min(cfl * case.chkdt(st["u"], st["v"], st["w"]), dtmax)
    for istep in range(1, 1501):
        case.rk3_step(st, dt)
        if istep % 10 == 0:
            dt = min(cfl * case.chkdt(st["u"], st["v"], st["w"]), dtmax)
            divtot, divmax = case.chkdiv(st["u"], st["v"], st["w"])
            assert divmax < 2.2e-9  # `small` in FP64, src/param.f90:18
    ref = np.loadtxt(os.path.join(HERE, "golden", "data_ldc_re1000.txt"))
    got = st["v"][1, 33, 1:65]  # v(1, 33, 1:64) in Fortran indices
    np.testing.assert_allclose(ref[:, 0], (np.arange(64) + 0.5) / 64, rtol=1e-7)
    np.testing.assert_allclose(got, ref[:, 1], rtol=1e-7, atol=0)
