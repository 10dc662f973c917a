"""The workloads at their stated sizes against the oracle (the C/OpenMP restatement of the
reference's FP64 path), on identical seeded inputs.

* Taylor-Green box 128^3, tri-periodic, side 2 pi, nu = 1/1600, dt = 1e-3: a Poisson solve
  of a seeded mean-free uniform[-1, 1] rhs, then one RK3 step from the Taylor-Green state
  with 1e-2 noise.  Bar: 1e-12 relative L-inf (pressure gauge-fixed: the tri-periodic
  problem is singular) and the divergence below the reference's `small` (param.f90:18).
* Heated cube 256^3: no-slip walls on all faces (velocity Dirichlet; the 3-D velocity
  Helmholtz systems take face-centred DST-I transforms along each component's normal),
  hot and cold x walls for the scalar (Dirichlet in x, Neumann in y and z), buoyancy,
  implicit diffusion of velocity and scalar (is_impdiff, main.f90:440-462), dt = 1e-3.
  Three RK3 steps in FP64 (1e-12) and in the mixed mode (1e-6, from an FP32-representable
  start).
"""
import time

import numpy as np
import pytest

from cans_amd import initflow as IF
from tests import cases

pytestmark = pytest.mark.gpu

SMALL = 2.220446049250313e-9


def relerr(a, b):
    return float(np.abs(a - b).max() / max(np.abs(b).max(), 1e-300))


def gauge(a):
    return a - a.mean()


@pytest.fixture(scope="module")
def torch():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    return torch


def test_tgv_128_poisson_and_step(torch):
    from cans_amd.host import Simulation
    p = cases.params("tgv")
    ng = p["ng"]
    assert ng == (128, 128, 128)
    print("\n[case] Taylor-Green %s: Poisson solve + one RK3 step" % (ng,), flush=True)
    sim = Simulation(cases.host_case(p))
    orc = cases.oracle_case(p)
    st = orc.new_state()
    # Poisson solve of a seeded mean-free rhs
    rhs = np.random.default_rng(128).uniform(-1, 1, ng)
    rhs -= rhs.mean()
    f = IF.halo(ng, rhs)
    sim.pp.set(f)
    sim.solver()
    sim.ctx.sync()
    st["pp"][...] = f
    orc.solver(st["pp"])
    err = relerr(gauge(sim.pp.interior()), gauge(st["pp"][1:-1, 1:-1, 1:-1]))
    print("poisson rel err (gauge-fixed) %.3e" % err, flush=True)
    assert err < 1e-12
    # one RK3 step
    u, v, w, pr = cases.initial_state("tgv", p, sim.grid, seed=3)
    for k, q in zip("uvwp", (u, v, w, pr)):
        st[k][...] = 0.
        st[k][1:-1, 1:-1, 1:-1] = q
    sim.set_state(*(IF.halo(ng, q) for q in (u, v, w, pr)))
    sim.init_bounds()
    orc.init_bounds(st)
    dt = 1e-3
    sim.rk3_step(dt)
    orc.rk3_step(st, dt)
    sim.ctx.sync()
    for k in "uvw":
        err = relerr(getattr(sim, k).numpy(), st[k])
        print("%s rel err %.3e" % (k, err), flush=True)
        assert err < 1e-12, (k, err)
    err = relerr(gauge(sim.p.interior()), gauge(st["p"][1:-1, 1:-1, 1:-1]))
    print("p rel err (gauge-fixed) %.3e" % err, flush=True)
    assert err < 1e-10
    _, mx = sim.chkdiv()
    _, omx = orc.chkdiv(st["u"], st["v"], st["w"])
    print("divmax gpu %.3e oracle %.3e" % (mx, omx), flush=True)
    assert mx < SMALL
    sim.close()


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_heated_cube_256_implicit_diffusion(torch, precision):
    from cans_amd import _lib as LB
    from cans_amd.host import Simulation
    p = cases.params("dhc3d")
    p["is_impdiff_1d"] = False
    p["is_impdiff"] = True
    ng = p["ng"]
    assert ng == (256, 256, 256)
    mixed = precision == "mixed"
    print("\n[case] heated cube %s, 3-D implicit diffusion, %s, 3 steps" % (ng, precision), flush=True)
    sim = Simulation(cases.host_case(p), precision=LB.CANS_MIXED if mixed else LB.CANS_FP64)
    u, v, w, pr = cases.initial_state("dhc3d", p, sim.grid, seed=9)
    if mixed:
        u, v, w, pr = (q.astype(np.float32).astype(np.float64) for q in (u, v, w, pr))
    sc = cases.initial_scalars("dhc3d", p, sim.grid, seed=10)
    sim.set_state(*(IF.halo(ng, q) for q in (u, v, w, pr)), s=[IF.halo(ng, q) for q in sc])
    sim.init_bounds()
    dt = 1e-3
    for _ in range(3):
        sim.rk3_step(dt)
    sim.ctx.sync()
    gpu = {k: getattr(sim, k).interior() for k in "uvwp"}
    gpu["s0"] = sim.s[0].interior()
    _, mx = sim.chkdiv()
    sim.close()
    del sim
    torch.cuda.empty_cache()
    orc = cases.oracle_case(p)
    st = orc.new_state()
    for k, q in zip("uvwp", (u, v, w, pr)):
        st[k][1:-1, 1:-1, 1:-1] = q
    st["s0"][1:-1, 1:-1, 1:-1] = sc[0]
    orc.init_bounds(st)
    for it in range(3):
        t0 = time.perf_counter()
        orc.rk3_step(st, dt)
        print("oracle step %d: %.1f s" % (it + 1, time.perf_counter() - t0), flush=True)
    tol = 1e-6 if mixed else 1e-12
    for k in ("u", "v", "w", "s0"):
        err = relerr(gpu[k], st[k][1:-1, 1:-1, 1:-1])
        print("%s rel err %.3e" % (k, err), flush=True)
        assert err < tol, (k, err)
    err = relerr(gauge(gpu["p"]), gauge(st["p"][1:-1, 1:-1, 1:-1]))
    print("p rel err (gauge-fixed) %.3e" % err, flush=True)
    assert err < (1e-6 if mixed else 1e-10)
    _, omx = orc.chkdiv(st["u"], st["v"], st["w"])
    print("divmax gpu %.3e oracle %.3e" % (mx, omx), flush=True)
    if not mixed:
        assert mx < SMALL
