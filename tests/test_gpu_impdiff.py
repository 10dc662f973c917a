"""Implicit diffusion on the GPU against the oracle (FP64 contexts).

The &numerics switches is_impdiff and is_impdiff_1d split the diffusion off the
explicit right-hand side (mom.f90:990-1016, scal.f90:86-97), build the Helmholtz
right-hand side in rk/rk_scal (rk.f90:312-326, 476-486), solve one Helmholtz
system per velocity component and scalar (main.f90:446-462; 3-D solver, or the
z-only elimination of solver_gaussel_z) and add alpha*lap(pp) in updatep.
* the z-only Helmholtz solve alone: bit for bit against the oracle;
* whole RK3 steps with scalars and buoyancy, fused and per-routine: 1e-12, including a
  heated cube with no-slip walls on all faces, whose 3-D velocity Helmholtz systems use
  the face-centred transforms (RODFT00 on the n-1 interior faces along the normal);
* the mixed mode (FP32 pressure side, FP64 velocity and Helmholtz solves): 1e-6 against
  the FP64 oracle per field, the pressure gauge-fixed when it is singular.
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


def imp_pair(name, ng, mode, precision=None, seed=7, f32_init=False):
    from cans_amd import _lib as LB
    from cans_amd.host import Simulation
    print("\n[case] %s ng=%s %s" % (name, ng, mode), flush=True)
    p = cases.params(name, ng=ng)
    p["is_impdiff"] = True
    p["is_impdiff_1d"] = mode == "1d"
    sim = Simulation(cases.host_case(p), precision=LB.CANS_FP64 if precision is None else precision)
    orc = cases.oracle_case(p)
    u, v, w, pr = cases.initial_state(name, p, sim.grid, seed=seed)
    sc = cases.initial_scalars(name, p, sim.grid, seed=seed + 1)
    if f32_init:  # both sides start from the same FP32-representable state
        u, v, w, pr = (q.astype(np.float32).astype(np.float64) for q in (u, v, w, pr))
    st = orc.new_state()
    for k, q in zip("uvwp", (u, v, w, pr)):
        st[k][1:-1, 1:-1, 1:-1] = q
    for i, q in enumerate(sc):
        st["s%d" % i][1:-1, 1:-1, 1:-1] = q
    sim.set_state(*(IF.halo(p["ng"], q) for q in (u, v, w, pr)), s=[IF.halo(p["ng"], q) for q in sc])
    return p, sim, orc, st


def test_helmholtz_z_bitwise(torch):
    p, sim, orc, st = imp_pair("rb_channel", (16, 8, 12), "1d")
    rng = np.random.default_rng(12)
    alpha = -0.37
    for k, (f, hp, ho) in enumerate(zip((sim.u, sim.v, sim.w), sim.hplans, orc.h_vel)):
        a = np.asfortranarray(rng.uniform(-1, 1, st["u"].shape))
        f.set(a)
        ref = a.copy(order="F")
        hp.solve_helmholtz(f, alpha)
        orc.solve_helmholtz(ref, alpha, ho)
        sim.ctx.sync()
        np.testing.assert_array_equal(f.numpy()[1:-1, 1:-1, 1:-1], ref[1:-1, 1:-1, 1:-1], err_msg="component %d" % k)


STEP_CASES = [("rb_channel", (32, 16, 24), "3d", 2e-3), ("rb_channel", (32, 16, 24), "1d", 2e-3),
              ("dhc", (64, 2, 64), "1d", 1e-2), ("dhc3d", (32, 24, 16), "3d", 2e-3)]


@pytest.mark.parametrize("name,ng,mode,dt", STEP_CASES, ids=["%s-%s" % (c[0], c[2]) for c in STEP_CASES])
@pytest.mark.parametrize("fused", [True, False], ids=["fused", "per-routine"])
def test_impdiff_steps_match_oracle(torch, name, ng, mode, dt, fused):
    p, sim, orc, st = imp_pair(name, ng, mode)
    sim.init_bounds()
    orc.init_bounds(st)
    for _ in range(3):
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
    dt_gpu, dt_orc = sim.chkdt(), orc.chkdt(st["u"], st["v"], st["w"])
    assert abs(dt_gpu - dt_orc) <= 1e-12 * dt_orc, (dt_gpu, dt_orc)


MIXED_CASES = [("rb_channel", (32, 16, 24), "3d", 2e-3), ("rb_channel", (32, 16, 24), "1d", 2e-3),
               ("dhc3d", (32, 24, 16), "3d", 2e-3)]


@pytest.mark.parametrize("name,ng,mode,dt", MIXED_CASES, ids=["%s-%s" % (c[0], c[2]) for c in MIXED_CASES])
@pytest.mark.parametrize("fused", [True, False], ids=["fused", "per-routine"])
def test_impdiff_mixed_steps(torch, name, ng, mode, dt, fused):
    from cans_amd import _lib as LB
    p, sim, orc, st = imp_pair(name, ng, mode, precision=LB.CANS_MIXED, f32_init=True)
    sim.init_bounds()
    orc.init_bounds(st)
    for _ in range(3):
        (sim.rk3_step if fused else sim.rk3_step_unfused)(dt)
        orc.rk3_step(st, dt)
    sim.ctx.sync()
    for k in "uvw":
        err = relerr(getattr(sim, k).interior(), st[k][1:-1, 1:-1, 1:-1])
        print("%s rel err %.3e" % (k, err), flush=True)
        assert err < MIXED_TOL, (k, err)
    for i in range(len(sim.s)):
        err = relerr(sim.s[i].interior(), st["s%d" % i][1:-1, 1:-1, 1:-1])
        print("s%d rel err %.3e" % (i, err), flush=True)
        assert err < MIXED_TOL, (i, err)
    a, b = sim.p.interior(), st["p"][1:-1, 1:-1, 1:-1]
    if "D" not in p["cbcpre"]:
        wgt = np.broadcast_to(1 / sim.grid.dzfi[1:-1], a.shape)
        a = a - (a * wgt).sum() / wgt.sum()
        b = b - (b * wgt).sum() / wgt.sum()
    err = relerr(a, b)
    print("p rel err (gauge-fixed) %.3e" % err, flush=True)
    assert err < MIXED_TOL, ("p", err)


@pytest.mark.parametrize("ng", [(32, 24, 16), (64, 48, 32)], ids=["32x24x16", "64x48x32"])
def test_walled_helmholtz_batch_bitwise(torch, ng):
    """The heated cube with no-slip walls and 3-D implicit diffusion: the fused substep solves the
    three velocity Helmholtz systems as one batch, one launch per pass although their x and y
    transforms differ (face-centred DST-I along each component's normal, cell-centred DST-II
    along the others); the per-routine sequence solves them one by one (cans_solve_helmholtz).
    Same arithmetic per line, so the same bits (FP64)."""
    from cans_amd.host import Simulation
    p, sim, orc, st = imp_pair("dhc3d", ng, "3d")
    sim2 = Simulation(cases.host_case(p))
    u, v, w, pr = cases.initial_state("dhc3d", p, sim.grid, seed=7)
    sc = cases.initial_scalars("dhc3d", p, sim.grid, seed=8)
    sim2.set_state(*(IF.halo(ng, q) for q in (u, v, w, pr)), s=[IF.halo(ng, q) for q in sc])
    sim.init_bounds()
    sim2.init_bounds()
    for _ in range(2):
        sim.rk3_step(2e-3)
        sim2.rk3_step_unfused(2e-3)
    sim.ctx.sync()
    sim2.ctx.sync()
    for k in ("u", "v", "w", "p"):
        np.testing.assert_array_equal(getattr(sim, k).interior(), getattr(sim2, k).interior(), err_msg=k)
    np.testing.assert_array_equal(sim.s[0].interior(), sim2.s[0].interior(), err_msg="s0")
