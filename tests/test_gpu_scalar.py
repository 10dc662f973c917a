"""Scalar transport and Boussinesq buoyancy on the GPU against the oracle.

* cans_rk_scal (scal.f90 + rk_scal): bit for bit (the bulk forcing f differs by
  the reduction order only, at the 1e-15 level);
* cans_rk with buoyancy: bit for bit;
* RK3 steps with two scalars (one forced, one with a source) and buoyancy, fused and
  per-routine, FP64 at 1e-12 and the mixed mode at 1e-6 against the oracle;
* the reference's differentially heated cavity (tests/differentially_heated_cavity):
  the GPU runs its 10000 adaptive-dt steps and must give the fixture's Nusselt number
  8.8252 to the reference's tolerance (rtol/atol 1e-2).
"""
import numpy as np
import pytest

from cans_amd import initflow as IF
from tests import cases

pytestmark = pytest.mark.gpu

RTOL = 1e-12
MIXED_TOL = 1e-6


def relerr(a, b):
    return float(np.abs(a - b).max() / max(np.abs(b).max(), 1e-300))


@pytest.fixture(scope="module")
def torch():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    return torch


def scalar_pair(name, ng, precision=None, seed=7):
    from cans_amd import _lib as LB
    from cans_amd.host import Simulation
    print("\n[case] %s ng=%s precision=%s" % (name, ng, precision), flush=True)
    p = cases.params(name, ng=ng)
    sim = Simulation(cases.host_case(p), precision=LB.CANS_FP64 if precision is None else precision)
    orc = cases.oracle_case(p)
    u, v, w, pr = cases.initial_state(name, p, sim.grid, seed=seed)
    sc = cases.initial_scalars(name, p, sim.grid, seed=seed + 1)
    if precision is not None and precision != LB.CANS_FP64:  # FP32-representable start for the mixed mode
        pr = pr.astype(np.float32).astype(np.float64)
    st = orc.new_state()
    for k, q in zip("uvwp", (u, v, w, pr)):
        st[k][1:-1, 1:-1, 1:-1] = q
    for i, q in enumerate(sc):
        st["s%d" % i][1:-1, 1:-1, 1:-1] = q
    sim.set_state(*(IF.halo(p["ng"], q) for q in (u, v, w, pr)), s=[IF.halo(p["ng"], q) for q in sc])
    return p, sim, orc, st


def test_rk_scal_and_buoyancy_bitwise(torch):
    from oracle.oracle import RKCOEFF
    p, sim, orc, st = scalar_pair("rb_channel", (16, 8, 12))
    rng = np.random.default_rng(4)
    for k in "uvwp":
        a = np.asfortranarray(rng.uniform(-1, 1, st[k].shape))
        st[k][...] = a
        getattr(sim, k).set(a)
    for i in range(2):
        a = np.asfortranarray(rng.uniform(-1, 1, st["s%d" % i].shape))
        st["s%d" % i][...] = a
        sim.s[i].set(a)
    dt = 2e-3
    for irk in (1, 2, 3):
        rkpar = RKCOEFF[irk - 1]
        for i, spec in enumerate(orc.scalars):
            f_gpu = sim.rk_scal(i, irk, dt)
            f_orc = orc.rk_scal(rkpar, dt, spec, st["u"], st["v"], st["w"], st["s%d" % i], st["dsdtrko%d" % i])
            sim.ctx.sync()
            np.testing.assert_array_equal(sim.s[i].numpy()[1:-1, 1:-1, 1:-1], st["s%d" % i][1:-1, 1:-1, 1:-1],
                                          err_msg="rk_scal irk=%d scalar %d" % (irk, i))
            np.testing.assert_array_equal(sim.dsdtrko[i].cpu().numpy(), st["dsdtrko%d" % i].ravel(order="F"))
            assert abs(f_gpu - f_orc) <= 1e-14 * max(1.0, abs(f_orc)), (f_gpu, f_orc)
            sim.bulk_forcing_s(i, f_orc)
            orc.bulk_forcing_s(spec["is_forced"], f_orc, st["s%d" % i])
            sim.boundp_s(i)
            orc.boundp(st["s%d" % i], cbc=spec["cbc"], bc=spec["bc"])
            sim.ctx.sync()
            np.testing.assert_array_equal(sim.s[i].numpy(), st["s%d" % i], err_msg="forcing+boundp %d" % i)
        sim.rk(irk, dt, with_f=False)
        orc.rk(rkpar, dt, st["p"], st["dudtrko"], st["dvdtrko"], st["dwdtrko"], st["u"], st["v"], st["w"],
               st["s0"])
        sim.ctx.sync()
        for k in "uvw":  # u is forced: f differs by the reduction order (1 ulp)
            err = relerr(getattr(sim, k).numpy()[1:-1, 1:-1, 1:-1], st[k][1:-1, 1:-1, 1:-1])
            assert err < 1e-14, (irk, k, err)
            getattr(sim, k).set(st[k])  # the next substep's scalar transport starts from equal velocities


STEP_CASES = [("rb_channel", (32, 16, 24), 1e-3, 3), ("dhc", (64, 2, 64), 5e-3, 3)]


@pytest.mark.parametrize("name,ng,dt,nsteps", STEP_CASES, ids=[c[0] for c in STEP_CASES])
@pytest.mark.parametrize("fused", [True, False], ids=["fused", "per-routine"])
def test_scalar_steps_match_oracle(torch, name, ng, dt, nsteps, fused):
    p, sim, orc, st = scalar_pair(name, ng)
    sim.init_bounds()
    orc.init_bounds(st)
    for _ in range(nsteps):
        (sim.rk3_step if fused else sim.rk3_step_unfused)(dt)
        orc.rk3_step(st, dt)
    sim.ctx.sync()
    for k in "uvw":
        err = relerr(getattr(sim, k).numpy(), st[k])
        print("%s rel err %.3e" % (k, err), flush=True)
        assert err < RTOL, (k, err)
    for i in range(len(sim.s)):
        err = relerr(sim.s[i].numpy(), st["s%d" % i])
        print("s%d rel err %.3e" % (i, err), flush=True)
        assert err < RTOL, (i, err)
    _, mx = sim.chkdiv()
    assert mx < 2.2e-9


def test_scalar_steps_mixed(torch):
    from cans_amd import _lib as LB
    p, sim, orc, st = scalar_pair("rb_channel", (32, 16, 24), precision=LB.CANS_MIXED)
    sim.init_bounds()
    orc.init_bounds(st)
    for _ in range(3):
        sim.rk3_step(1e-3)
        orc.rk3_step(st, 1e-3)
    sim.ctx.sync()
    for k in "uvw":
        err = relerr(getattr(sim, k).numpy(), st[k])
        print("mixed %s rel err %.3e" % (k, err), flush=True)
        assert err < MIXED_TOL, (k, err)
    for i in range(len(sim.s)):
        err = relerr(sim.s[i].numpy(), st["s%d" % i])
        print("mixed s%d rel err %.3e" % (i, err), flush=True)
        assert err < MIXED_TOL, (i, err)


def test_dhc_nusselt_fixture_gpu(torch):
    """tests/differentially_heated_cavity on the GPU: 10000 fused steps, dt from chkdt
    every icheck = 10 steps (cfl 0.95, dtmax 1e9), Nusselt of test.py:7-20."""
    p = cases.params("dhc")
    from cans_amd.host import Simulation
    sim = Simulation(cases.host_case(p))
    sc = cases.initial_scalars("dhc", p, sim.grid)
    zero = np.zeros(tuple(p["ng"]))
    sim.set_state(*(IF.halo(p["ng"], zero) for _ in range(4)), s=[IF.halo(p["ng"], sc[0])])
    sim.init_bounds()
    cfl, dtmax = 0.95, 1e9
    dt = min(cfl * sim.chkdt(), dtmax)
    for istep in range(1, 10001):
        sim.rk3_step(dt)
        if istep % 10 == 0:
            dt = min(cfl * sim.chkdt(), dtmax)
        if istep % 2000 == 0:
            print("[dhc] step %d dt %.4e" % (istep, dt), flush=True)
    _, mx = sim.chkdiv()
    nu = cases.dhc_nusselt(sim.s[0].numpy()[1:-1, 1:-1, 1:-1], sim.grid.dl)
    print("[dhc] Nusselt %.5f (fixture 8.8252), divmax %.3e" % (nu, mx), flush=True)
    assert mx < 2.2e-9
    np.testing.assert_allclose([nu], [8.8252], rtol=1e-2, atol=1e-2)
