"""Transform lengths the FFT passes do not take (primes, 2^m with m < 3 times odd factors,
more than three odd factors, odd n1): cans_fftini gives such a direction the dense matrix of
its kind (cans_transform_matrix, pinned against the oracle on the CPU), so the solver takes
any length, as FFTW does (fft.f90:87-97).  Against the oracle: Poisson solves at the FP64 bar
(1e-12; singular systems gauge-fixed) and the mixed mode's 1e-6, face-centred Helmholtz
solves with walls, and whole RK3 steps (the divergence then comes from fillps, as the
reference computes it, instead of the first FFT pass).
"""
import numpy as np
import pytest

from cans_amd import initflow as IF
from tests import cases

pytestmark = pytest.mark.gpu


def relerr(a, b):
    return float(np.abs(a - b).max() / max(np.abs(b).max(), 1e-300))


def gauge(a, dzf):
    wgt = np.broadcast_to(dzf, a.shape)
    return a - (a * wgt).sum() / wgt.sum()


@pytest.fixture(scope="module")
def torch():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    return torch


def pair(name, ng, precision="fp64", p=None, seed=7):
    from cans_amd import _lib as LB
    from cans_amd.host import Simulation
    print("\n[case] %s ng=%s %s" % (name, ng, precision), flush=True)
    p = p or cases.params(name, ng=ng)
    mixed = precision == "mixed"
    sim = Simulation(cases.host_case(p), precision=LB.CANS_MIXED if mixed else LB.CANS_FP64)
    orc = cases.oracle_case(p)
    init = cases.initial_state(name, p, sim.grid, seed=seed)
    if mixed:
        init = [q.astype(np.float32).astype(np.float64) for q in init]
    st = orc.new_state()
    for k, q in zip("uvwp", init):
        st[k][1:-1, 1:-1, 1:-1] = q
    sim.set_state(*(IF.halo(p["ng"], q) for q in init))
    return p, sim, orc, st


# x 14 = 2*7, x 12 = 4*3 (m < 3), y 270 = 2*27*5 (four odd factors), odd n1 (both x and y dense),
# DST/DCT and REDFT11/RODFT11 kinds at prime lengths, a periodic z
SOLVER_CASES = [("channel", (14, 16, 12)), ("channel", (12, 16, 10)), ("channel", (16, 270, 8)),
                ("channel", (45, 20, 12)), ("cavity_dd", (21, 14, 10)), ("box", (13, 11, 9)),
                ("open_nd", (7, 11, 10)), ("open_x", (19, 16, 8)), ("tgv", (15, 14, 12))]


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
@pytest.mark.parametrize("name,ng", SOLVER_CASES, ids=["%s-%dx%dx%d" % ((c,) + g) for c, g in SOLVER_CASES])
def test_dense_solver_matches_oracle(torch, name, ng, precision):
    p, sim, orc, st = pair(name, ng, precision)
    rng = np.random.default_rng(11)
    rhs = rng.uniform(-1, 1, tuple(ng))
    dzf = 1 / sim.grid.dzfi[1:-1]
    singular = "D" not in p["cbcpre"]
    if singular:
        rhs = gauge(rhs, dzf)
    if precision == "mixed":
        rhs = rhs.astype(np.float32).astype(np.float64)
    f = IF.halo(ng, rhs)
    sim.pp.set(f)
    sim.solver()
    sim.ctx.sync()
    st["pp"][...] = f
    orc.solver(st["pp"])
    got, ref = sim.pp.interior(), st["pp"][1:-1, 1:-1, 1:-1]
    if singular:
        got, ref = gauge(got, dzf), gauge(ref, dzf)
    err = relerr(got, ref)
    print("dense solver rel err %.3e" % err, flush=True)
    assert err < (1e-6 if precision == "mixed" else 1e-12)


WALLS = [("fcc", "DDPPDD", (0.3, -0.2, 0.0, 0.0, 0.5, -0.7), (21, 14, 12)),
         ("cfc", "PPDDNN", (0.0, 0.0, -0.6, 0.35, 0.2, -0.15), (18, 13, 10)),
         ("fcc", "DNDDDD", (0.2, 0.1, 0.3, -0.3, 0.1, 0.4), (11, 12, 10)),
         ("fcc", "NDNNDD", (0.2, 0.1, 0.3, -0.3, 0.1, 0.4), (9, 14, 8))]


@pytest.mark.parametrize("c_or_f,cbc,bc,ng", WALLS, ids=["%s-%s" % w[:2] for w in WALLS])
def test_dense_helmholtz_face_centred_walls(torch, c_or_f, cbc, bc, ng):
    from cans_amd.host import Plan, Simulation
    print("\n[case] dense helmholtz walls %s %s ng=%s" % (c_or_f, cbc, ng), flush=True)
    pre = "".join("P" if t == "P" else "N" for t in cbc)
    p = dict(ng=ng, l=(1.1, 0.8, 1.0), gtype=1, gr=1.5, cbcvel=pre * 3, cbcpre=pre, visci=60.0)
    sim = Simulation(cases.host_case(p))
    orc = cases.oracle_case(p)
    plan = Plan(sim.ctx, sim.grid, cbc, bc, c_or_f)
    s = orc.initsolver(cbc, np.asarray(bc), c_or_f)
    alpha = -0.013
    rng = np.random.default_rng(8)
    rhs = IF.halo(ng, rng.uniform(-1, 1, tuple(ng)))
    f = sim.ctx.field()
    f.set(rhs)
    plan.solve_helmholtz(f, alpha)
    sim.ctx.sync()
    ref = rhs.copy(order="F")
    orc.solve_helmholtz(ref, alpha, s)
    q = 1 if (c_or_f[2] == "f" and cbc[5] == "D") else 0  # the last z face is not solved (solver.f90:77)
    err = relerr(f.interior()[:, :, :ng[2] - q], ref[1:-1, 1:-1, 1:-1][:, :, :ng[2] - q])
    print("dense helmholtz rel err %.3e" % err, flush=True)
    plan.close()
    assert err < 1e-12


STEP_CASES = [("channel", (14, 12, 16), 2e-4, "fp64"), ("channel", (14, 12, 16), 2e-4, "mixed"),
              ("tgv", (15, 15, 15), 1e-3, "fp64"), ("duct", (18, 13, 12), 1e-3, "fp64")]


@pytest.mark.parametrize("name,ng,dt,precision", STEP_CASES, ids=["%s-%dx%dx%d-%s" % ((c[0],) + c[1] + (c[3],))
                                                                  for c in STEP_CASES])
def test_dense_steps_match_oracle(torch, name, ng, dt, precision):
    p, sim, orc, st = pair(name, ng, precision)
    sim.init_bounds()
    orc.init_bounds(st)
    for _ in range(3):
        sim.rk3_step(dt)
        orc.rk3_step(st, dt)
    sim.ctx.sync()
    tol = 1e-6 if precision == "mixed" else 1e-12
    for k in "uvw":
        err = relerr(getattr(sim, k).interior(), st[k][1:-1, 1:-1, 1:-1])
        print("%s rel err %.3e" % (k, err), flush=True)
        assert err < tol, (k, err)
    dzf = 1 / sim.grid.dzfi[1:-1]
    got, ref = sim.p.interior(), st["p"][1:-1, 1:-1, 1:-1]
    if "D" not in p["cbcpre"]:
        got, ref = gauge(got, dzf), gauge(ref, dzf)
    err = relerr(got, ref)
    print("p rel err %.3e" % err, flush=True)
    assert err < (1e-6 if precision == "mixed" else 1e-9)
