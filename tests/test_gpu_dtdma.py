"""The distributed tridiagonal z solve (gaussel_dtdma, solver.f90:309-517) on the GPU.

With the z rows split into slabs (the ranks of a z decomposition), the pressure
solve must give the solution of the plain sweep: against the oracle's plain
gaussel at the FP64 bar, for slab counts that do and do not divide n3, with slabs
of 3 rows (no interior block), of fewer rows than one checkpoint block and of
several blocks.  A periodic z makes the reduced system cyclic (solver.f90:469-498):
the same bar against the oracle's periodic gaussel.  Whole RK3 steps of the
stretched channel and of the tri-periodic Taylor-Green box then match the oracle as
with the plain sweep (velocity 1e-12; the pressure's zero mode is compared
gauge-fixed), and with FP32 storage at the mixed mode's 1e-6 bar.
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


def dt_pair(name, ng, nslab, seed=7, precision="fp64", p=None):
    from cans_amd import _lib as LB
    from cans_amd.host import Simulation
    print("\n[case] %s ng=%s slabs=%d %s" % (name, ng, nslab, precision), flush=True)
    p = p or cases.params(name, ng=ng)
    p["dtdma_slabs"] = nslab
    sim = Simulation(cases.host_case(p), precision=LB.CANS_MIXED if precision == "mixed" else LB.CANS_FP64)
    orc = cases.oracle_case(p)
    u, v, w, pr = cases.initial_state(name, p, sim.grid, seed=seed)
    if precision == "mixed":
        u, v, w, pr = (q.astype(np.float32).astype(np.float64) for q in (u, v, w, pr))
    st = orc.new_state()
    for k, q in zip("uvwp", (u, v, w, pr)):
        st[k][1:-1, 1:-1, 1:-1] = q
    sim.set_state(*(IF.halo(p["ng"], q) for q in (u, v, w, pr)))
    return p, sim, orc, st


def gauge(a, dzf):
    wgt = np.broadcast_to(dzf, a.shape)
    return a - (a * wgt).sum() / wgt.sum()


# (n3, slabs): 24/2 = 12-row slabs (two checkpoint blocks), 24/3, 24/5 (slabs of 5 and 4 rows),
# 24/8 (3-row slabs: no interior block beyond one row), 61/4 (16 and 15 rows), 120/3 (40 rows: five blocks)
SOLVER_SLABS = [(24, 2), (24, 3), (24, 5), (24, 8), (61, 4), (120, 3)]


@pytest.mark.parametrize("n3,nslab", SOLVER_SLABS)
def test_dtdma_solver_matches_oracle(torch, n3, nslab):
    # Dirichlet pressure walls in y: no zero wavenumber, every line regular
    p, sim, orc, st = dt_pair("cavity_dd", (16, 32, n3), nslab)
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


def periodic_z_case(ng):
    """Outflow (pressure Dirichlet) faces at both x ends, periodic y and z, uniform z: every
    line regular, so the cyclic reduced system is compared without a gauge."""
    return dict(ng=ng, l=(1.0, 1.0, 1.0), gtype=1, gr=0.0, cbcvel="NNPPPP" * 3, cbcpre="DDPPPP", visci=100.0)


@pytest.mark.parametrize("n3,nslab", [(24, 2), (24, 3), (25, 4), (48, 8), (61, 5)])
def test_dtdma_periodic_solver_matches_oracle(torch, n3, nslab):
    p, sim, orc, st = dt_pair("periodic_z", (16, 24, n3), nslab, p=periodic_z_case((16, 24, n3)))
    rng = np.random.default_rng(12)
    rhs = rng.uniform(-1, 1, tuple(p["ng"]))
    f = IF.halo(p["ng"], rhs)
    sim.pp.set(f)
    sim.solver()
    sim.ctx.sync()
    st["pp"][...] = f
    orc.solver(st["pp"])
    err = relerr(sim.pp.numpy()[1:-1, 1:-1, 1:-1], st["pp"][1:-1, 1:-1, 1:-1])
    print("periodic dtdma solver rel err %.3e" % err, flush=True)
    assert err < RTOL


def test_dtdma_periodic_tgv_steps_match_oracle(torch):
    """Tri-periodic: the zero-wavenumber line is singular; its last unknown is pinned, so p is
    compared gauge-fixed and the velocity must match as with the plain sweep."""
    p, sim, orc, st = dt_pair("tgv", (32, 32, 30), 3)
    sim.init_bounds()
    orc.init_bounds(st)
    for _ in range(3):
        sim.rk3_step(1e-3)
        orc.rk3_step(st, 1e-3)
    sim.ctx.sync()
    for k in "uvw":
        err = relerr(getattr(sim, k).numpy(), st[k])
        print("%s rel err %.3e" % (k, err), flush=True)
        assert err < RTOL, (k, err)
    dzf = 1 / sim.grid.dzfi[1:-1]
    err = relerr(gauge(sim.p.numpy()[1:-1, 1:-1, 1:-1], dzf), gauge(st["p"][1:-1, 1:-1, 1:-1], dzf))
    print("p rel err (gauge-fixed) %.3e" % err, flush=True)
    assert err < 1e-9
    _, mx = sim.chkdiv()
    assert mx < 2.2e-9


@pytest.mark.parametrize("name,ng,nslab,dt", [("channel", (32, 16, 48), 4, 2e-4), ("tgv", (32, 32, 30), 3, 1e-3)])
def test_dtdma_mixed_steps(torch, name, ng, nslab, dt):
    """FP32 storage: the singular line's rhs is compatible only to FP32 rounding; with its last
    unknown pinned the solve stays at the mixed mode's bar (unpinned, the constant mode would
    swamp the FP32 pressure)."""
    p, sim, orc, st = dt_pair(name, ng, nslab, precision="mixed")
    sim.init_bounds()
    orc.init_bounds(st)
    for _ in range(5):
        sim.rk3_step(dt)
        orc.rk3_step(st, dt)
    sim.ctx.sync()
    for k in "uvw":
        err = relerr(getattr(sim, k).interior(), st[k][1:-1, 1:-1, 1:-1])
        print("%s rel err %.3e" % (k, err), flush=True)
        assert err < 1e-6, (k, err)
    dzf = 1 / sim.grid.dzfi[1:-1]
    err = relerr(gauge(sim.p.interior(), dzf), gauge(st["p"][1:-1, 1:-1, 1:-1], dzf))
    print("p rel err (gauge-fixed) %.3e" % err, flush=True)
    assert err < 1e-6


def test_dtdma_thin_slabs_refused(torch):
    from cans_amd import _lib as LB
    with pytest.raises(LB.CansError):
        p, sim, orc, st = dt_pair("cavity_dd", (16, 32, 24), 9)
        sim.pp.set(IF.halo(p["ng"], np.zeros(tuple(p["ng"]))))
        sim.solver()
