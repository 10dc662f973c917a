"""The distributed tridiagonal z solve (gaussel_dtdma, solver.f90:309-517) on the GPU.

With the z rows split into slabs (the ranks of a z decomposition), the pressure
solve must give the solution of the plain sweep: against the oracle's plain
gaussel at the FP64 bar, for slab counts that do and do not divide n3.  Whole RK3
steps of the stretched channel then match the oracle as with the plain sweep
(velocity 1e-12; the pressure's zero mode, unpinned in the reference's
gaussel_dtdma too, is compared gauge-fixed).
"""
import numpy as np
import pytest

from cans_amd import initflow as IF
from tests import cases

pytestmark = pytest.mark.gpu

RTOL = 1e-12


def relerr(a, b):
    return float(np.abs(a - b).max() / max(np.abs(b).max(), 1e-300))


@pytest.fixture(scope="module")
def torch():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    return torch


def dt_pair(name, ng, nslab, seed=7):
    from cans_amd.host import Simulation
    print("\n[case] %s ng=%s slabs=%d" % (name, ng, nslab), flush=True)
    p = cases.params(name, ng=ng)
    p["dtdma_slabs"] = nslab
    sim = Simulation(cases.host_case(p))
    orc = cases.oracle_case(p)
    u, v, w, pr = cases.initial_state(name, p, sim.grid, seed=seed)
    st = orc.new_state()
    for k, q in zip("uvwp", (u, v, w, pr)):
        st[k][1:-1, 1:-1, 1:-1] = q
    sim.set_state(*(IF.halo(p["ng"], q) for q in (u, v, w, pr)))
    return p, sim, orc, st


@pytest.mark.parametrize("nslab", [2, 3, 5])
def test_dtdma_solver_matches_oracle(torch, nslab):
    # Dirichlet pressure walls in y: no zero wavenumber, every line regular
    p, sim, orc, st = dt_pair("cavity_dd", (16, 32, 24), nslab)
    rng = np.random.default_rng(11)
    rhs = rng.uniform(-1, 1, tuple(p["ng"]))
    f = IF.halo(p["ng"], rhs)
    sim.pp.set(f)
    sim.solver()
    sim.ctx.sync()
    st["pp"][...] = f
    orc.solver(st["pp"])
    err = relerr(sim.pp.numpy()[1:-1, 1:-1, 1:-1], st["pp"][1:-1, 1:-1, 1:-1])
    print("dtdma solver rel err %.3e" % err, flush=True)
    assert err < RTOL


def test_dtdma_steps_match_oracle(torch):
    p, sim, orc, st = dt_pair("channel", (32, 16, 24), 4)
    sim.init_bounds()
    orc.init_bounds(st)
    for _ in range(3):
        sim.rk3_step(2e-4)
        orc.rk3_step(st, 2e-4)
    sim.ctx.sync()
    for k in "uvw":
        err = relerr(getattr(sim, k).numpy(), st[k])
        print("%s rel err %.3e" % (k, err), flush=True)
        assert err < RTOL, (k, err)
    a = sim.p.numpy()[1:-1, 1:-1, 1:-1]
    b = st["p"][1:-1, 1:-1, 1:-1]
    wgt = np.broadcast_to(1 / sim.grid.dzfi[1:-1], a.shape)
    err = relerr(a - (a * wgt).sum() / wgt.sum(), b - (b * wgt).sum() / wgt.sum())
    print("p rel err (gauge-fixed) %.3e" % err, flush=True)
    assert err < 1e-9
    _, mx = sim.chkdiv()
    assert mx < 2.2e-9
