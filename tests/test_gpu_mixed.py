"""Mixed precision (CANS_MIXED: velocity stored as FP64; p, pp, RHS_old, the RK
increment and the transform passes stored as FP32; all arithmetic in FP64)
against the FP64 oracle on identical inputs.

Bar: 1e-6 relative L-inf per field (max |mixed - oracle| / max |oracle|) for u,
v, w and the gauge-fixed pressure after 10 RK3 steps, and for single Poisson
solves.  The divergence left after the projection comes from the FP32-stored
pressure correction; it is held to 1e-6 of the velocity-gradient scale
max|u| dxi + max|v| dyi + max|w| max(dzfi).
"""
import numpy as np
import pytest

from cans_amd import initflow as IF
from tests import cases

pytestmark = pytest.mark.gpu

MIXED_TOL = 1e-6


def relerr(a, b):
    return float(np.abs(a - b).max() / max(np.abs(b).max(), 1e-300))


def gauge_fixed(a, dzf, singular):
    if not singular:
        return a
    wgt = np.broadcast_to(dzf, a.shape)
    return a - (a * wgt).sum() / wgt.sum()


@pytest.fixture(scope="module")
def torch():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    return torch


def mixed_pair(name, ng, seed=7):
    from cans_amd import _lib as LB
    from cans_amd.host import Simulation
    print("\n[case] mixed %s ng=%s seed=%d" % (name, ng, seed), flush=True)
    p = cases.params(name, ng=ng)
    sim = Simulation(cases.host_case(p), precision=LB.CANS_MIXED)
    orc = cases.oracle_case(p)
    init = [q.astype(np.float32).astype(np.float64) for q in cases.initial_state(name, p, sim.grid, seed=seed)]
    st = orc.new_state()
    for k, q in zip("uvwp", init):
        st[k][1:-1, 1:-1, 1:-1] = q
    sim.set_state(*(IF.halo(p["ng"], q) for q in init))
    return p, sim, orc, st


SOLVER_CASES = [("tgv", (32, 32, 32)), ("channel", (64, 32, 48)), ("duct", (16, 32, 20)),
                ("cavity_dd", (16, 32, 24)), ("channel", (128, 64, 96)), ("channel", (96, 48, 40)),
                ("channel", (160, 120, 40)), ("open_nd", (64, 32, 20)), ("open_x", (128, 16, 12))]


@pytest.mark.parametrize("name,ng", SOLVER_CASES, ids=["%s-%dx%dx%d" % ((c,) + g) for c, g in SOLVER_CASES])
def test_mixed_solver(torch, name, ng):
    p, sim, orc, st = mixed_pair(name, ng)
    rng = np.random.default_rng(11)
    rhs = rng.uniform(-1, 1, tuple(ng))
    singular = "D" not in p["cbcpre"]
    dzf = 1 / sim.grid.dzfi[1:-1]
    if singular:
        rhs = gauge_fixed(rhs, dzf, True)
    rhs = rhs.astype(np.float32).astype(np.float64)
    f = IF.halo(ng, rhs)
    sim.pp.set(f)
    sim.solver()
    sim.ctx.sync()
    st["pp"][...] = f
    orc.solver(st["pp"])
    a = gauge_fixed(sim.pp.interior(), dzf, singular)
    b = gauge_fixed(st["pp"][1:-1, 1:-1, 1:-1], dzf, singular)
    err = relerr(a, b)
    print("mixed solver rel err (gauge-fixed) %.3e" % err, flush=True)
    assert err < MIXED_TOL


# the walled box (no periodic x) takes the unfused divergence path (fillps, then the solve)
STEP_CASES = [("channel", (64, 32, 48), 2e-4, 10), ("tgv", (32, 32, 32), 1e-3, 10), ("duct", (32, 16, 24), 1e-3, 5),
              ("box", (16, 16, 12), 1e-3, 5), ("open_nd", (16, 24, 12), 1e-3, 5), ("channel", (80, 40, 24), 2e-4, 5)]


# the intermediate velocity as an FP64 field (the default) and as FP32 increments
# (cans_ctx_set_mixed_increments) for the fused step; the per-routine sequence has no such choice
@pytest.mark.parametrize("name,ng,dt,nsteps", STEP_CASES, ids=["%s-%dx%dx%d" % ((c[0],) + c[1]) for c in STEP_CASES])
@pytest.mark.parametrize("fused", ["fused", "fused-increments", "per-routine"])
def test_mixed_steps(torch, name, ng, dt, nsteps, fused):
    p, sim, orc, st = mixed_pair(name, ng)
    if fused == "fused-increments":
        sim.set_mixed_increments(True)
    fused = fused != "per-routine"
    sim.init_bounds()
    orc.init_bounds(st)
    for it in range(nsteps):
        (sim.rk3_step if fused else sim.rk3_step_unfused)(dt)
        orc.rk3_step(st, dt)
    sim.ctx.sync()
    for k in "uvw":
        err = relerr(getattr(sim, k).interior(), st[k][1:-1, 1:-1, 1:-1])
        print("%s rel err %.3e" % (k, err), flush=True)
        assert err < MIXED_TOL, (k, err)
    singular = "D" not in p["cbcpre"]
    dzf = 1 / sim.grid.dzfi[1:-1]
    err = relerr(gauge_fixed(sim.p.interior(), dzf, singular), gauge_fixed(st["p"][1:-1, 1:-1, 1:-1], dzf, singular))
    print("p rel err (gauge-fixed) %.3e" % err, flush=True)
    assert err < MIXED_TOL, ("p", err)
    tot, mx = sim.chkdiv()
    otot, omx = orc.chkdiv(st["u"], st["v"], st["w"])
    g = sim.grid
    vmax = [np.abs(getattr(sim, k).interior()).max() for k in "uvw"]
    scale = vmax[0] * g.dli[0] + vmax[1] * g.dli[1] + vmax[2] * g.dzfi.max()
    print("divmax mixed %.3e oracle %.3e (relative to the gradient scale %.3e)" % (mx, omx, mx / scale), flush=True)
    assert mx <= 1e-6 * scale


# FP64 refinement of the mixed pressure solve (cans_ctx_set_refinement): every step's velocity
# must come out divergence-free to the reference's FP64 abort threshold `small` (param.f90:18,
# main.f90:501-510), not only to FP32 level.  mask 4 refines the last substep of each step,
# whose projection alone sets the divergence the step leaves; mask 7 refines every substep.
SMALL = 2.220446049250313e-9
REFINED = [("channel", (64, 32, 48), 2e-4, 5, 2, 4), ("channel", (64, 32, 48), 2e-4, 5, 1, 7),
           ("tgv", (32, 32, 32), 1e-3, 5, 2, 7), ("duct", (32, 16, 24), 1e-3, 4, 2, 4),
           ("box", (16, 16, 12), 1e-3, 4, 2, 4), ("open_nd", (16, 24, 12), 1e-3, 4, 2, 7)]


@pytest.mark.parametrize("increments", [False, True], ids=["fp64-intermediate", "increments"])
@pytest.mark.parametrize("name,ng,dt,nsteps,sweeps,mask", REFINED,
                         ids=["%s-%dx%dx%d-s%d-m%d" % ((c[0],) + c[1] + c[4:]) for c in REFINED])
def test_mixed_refined_steps(torch, name, ng, dt, nsteps, sweeps, mask, increments):
    p, sim, orc, st = mixed_pair(name, ng)
    sim.set_refinement(sweeps, mask)
    if increments:
        sim.set_mixed_increments(True)
    sim.init_bounds()
    orc.init_bounds(st)
    g = sim.grid
    for it in range(nsteps):
        sim.rk3_step(dt)
        orc.rk3_step(st, dt)
        tot, mx = sim.chkdiv()
        vmax = [np.abs(getattr(sim, k).interior()).max() for k in "uvw"]
        scale = vmax[0] * g.dli[0] + vmax[1] * g.dli[1] + vmax[2] * g.dzfi.max()
        print("step %d divmax %.3e (%.2e of the gradient scale)" % (it + 1, mx, mx / scale), flush=True)
        assert mx < SMALL
        assert mx <= 1e-12 * scale
    for k in "uvw":
        err = relerr(getattr(sim, k).interior(), st[k][1:-1, 1:-1, 1:-1])
        print("%s rel err %.3e" % (k, err), flush=True)
        assert err < MIXED_TOL, (k, err)
    singular = "D" not in p["cbcpre"]
    dzf = 1 / g.dzfi[1:-1]
    err = relerr(gauge_fixed(sim.p.interior(), dzf, singular), gauge_fixed(st["p"][1:-1, 1:-1, 1:-1], dzf, singular))
    print("p rel err (gauge-fixed) %.3e" % err, flush=True)
    assert err < MIXED_TOL, ("p", err)


def test_increments_setting_refused_outside_mixed(torch):
    from cans_amd import _lib as LB
    from cans_amd.host import Simulation
    sim = Simulation(cases.host_case(cases.params("tgv", ng=(16, 16, 16))))
    with pytest.raises(LB.CansError):
        sim.set_mixed_increments(True)
    sim.set_mixed_increments(False)


def test_refinement_setting_refused_outside_mixed(torch):
    from cans_amd import _lib as LB
    from cans_amd.host import Simulation
    p = cases.params("tgv", ng=(16, 16, 16))
    sim = Simulation(cases.host_case(p))
    with pytest.raises(LB.CansError):
        sim.set_refinement(2, 4)
    sim.set_refinement(0, 0)
    sim.close()
