"""Parity at the benchmark's full size: the 512x256x192 stretched periodic channel.

Ten RK3 steps at dt = 2e-4 on the GPU and in the oracle from the same seeded state (velocity
at 1e-12 relative L-inf, pressure gauge-fixed within 4x the oracle's own FFT
spread as in test_gpu_parity), plus size-independent properties: the projected
velocity is divergence-free to the reference's FP64 `small`, and the discrete
Laplacian of the Poisson solution reproduces its right-hand side.
"""
import numpy as np
import pytest

from cans_amd import initflow as IF
from tests import cases

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def setup():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    from cans_amd.host import Simulation
    p = cases.params("channel")
    print("\n[case] channel full size %s" % (p["ng"],), flush=True)
    sim = Simulation(cases.host_case(p))
    u, v, w, pr = cases.initial_state("channel", p, sim.grid, seed=1)
    return p, sim, (u, v, w, pr)


def test_full_size_step_matches_oracle(setup):
    from oracle import oracle as O
    p, sim, init = setup
    sim.set_state(*(IF.halo(p["ng"], q) for q in init))
    sim.init_bounds()
    dt = 2e-4
    nsteps = 10
    for _ in range(nsteps):
        sim.rk3_step(dt)
    sim.ctx.sync()
    orc = cases.oracle_case(p)
    st = orc.new_state()
    for k, q in zip("uvwp", init):
        st[k][1:-1, 1:-1, 1:-1] = q
    orc.init_bounds(st)
    for _ in range(nsteps):
        orc.rk3_step(st, dt)
    for k in "uvw":
        a, b = getattr(sim, k).numpy(), st[k]
        err = np.abs(a - b).max() / np.abs(b).max()
        print("%s rel err %.3e" % (k, err), flush=True)
        assert err < 1e-12
    a = sim.p.interior()
    b = st["p"][1:-1, 1:-1, 1:-1]
    wgt = np.broadcast_to(1 / sim.grid.dzfi[1:-1], a.shape)
    da = a - (a * wgt).sum() / wgt.sum()
    db = b - (b * wgt).sum() / wgt.sum()
    err = np.abs(da - db).max() / np.abs(db).max()
    print("p rel err (gauge-fixed) %.3e" % err, flush=True)
    assert err < 1e-10
    tot, mx = sim.chkdiv()
    print("divmax %.3e" % mx, flush=True)
    assert mx < 2.2e-9


def test_full_size_poisson_residual(setup):
    p, sim, init = setup
    rng = np.random.default_rng(4)
    ng = p["ng"]
    rhs = rng.uniform(-1, 1, tuple(ng))
    wgt = np.broadcast_to(1 / sim.grid.dzfi[1:-1], rhs.shape)
    rhs -= (rhs * wgt).sum() / wgt.sum()
    sim.pp.set(IF.halo(ng, rhs))
    sim.solver()
    sim.boundp(sim.pp)
    sim.ctx.sync()
    q = sim.pp.numpy()
    g = sim.grid
    k = np.arange(1, ng[2] + 1)
    c = q[1:-1, 1:-1, 1:-1]
    lap = ((q[2:, 1:-1, 1:-1] - 2 * c + q[:-2, 1:-1, 1:-1]) * g.dli[0] ** 2 +
           (q[1:-1, 2:, 1:-1] - 2 * c + q[1:-1, :-2, 1:-1]) * g.dli[1] ** 2 +
           ((q[1:-1, 1:-1, 2:] - c) * g.dzci[k] - (c - q[1:-1, 1:-1, :-2]) * g.dzci[k - 1]) * g.dzfi[k])
    res = np.abs(lap - rhs).max()
    print("poisson residual %.3e" % res, flush=True)
    assert res < 1e-9 * np.abs(rhs).max()
