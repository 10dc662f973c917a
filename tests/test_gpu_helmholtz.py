"""Implicit-diffusion Helmholtz solves (solve_helmholtz.f90:28-75) on the GPU against the oracle.

The three velocity systems of a stretched Couette channel: (1 + alpha*lap) q = rhs with
c_or_f = fcc, cfc, ccf for u, v, w (main.f90:457-462), periodic x and y, no-slip walls in z
with the upper wall moving in x, so that the boundary rhs (initsolver.f90:189-232 through
updt_rhs_b, bound.f90:514-598) is non-zero for u.  Bar: 1e-12 relative L-inf (FP64).
"""
import numpy as np
import pytest

from cans_amd import initflow as IF
from tests import cases

pytestmark = pytest.mark.gpu


def component_cbc(cbcvel, cc):
    return "".join(cbcvel[ib + 2 * d + 6 * cc] for d in range(3) for ib in range(2))


def component_bc(bcvel, cc):
    return [bcvel[ib + 2 * d + 6 * cc] for d in range(3) for ib in range(2)]


CASES = [((64, 32, 48), -1e-3), ((32, 16, 24), -0.05)]


@pytest.mark.parametrize("ng,alpha", CASES, ids=["%dx%dx%d" % c[0] for c in CASES])
@pytest.mark.parametrize("cc,c_or_f", [(0, "fcc"), (1, "cfc"), (2, "ccf")], ids=["u", "v", "w"])
def test_helmholtz_matches_oracle(ng, alpha, cc, c_or_f):
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    from cans_amd.host import Plan, Simulation
    print("\n[case] helmholtz channel ng=%s alpha=%g component %d (%s)" % (ng, alpha, cc, c_or_f), flush=True)
    p = cases.params("channel", ng=ng)
    bcvel = np.zeros(18)
    bcvel[1 + 4 + 0] = 1.0  # upper z wall moves in x (u)
    p["bcvel"] = tuple(bcvel)
    sim = Simulation(cases.host_case(p))
    orc = cases.oracle_case(p)
    cbc = component_cbc(p["cbcvel"], cc)
    bc = component_bc(bcvel, cc)
    plan = Plan(sim.ctx, sim.grid, cbc, bc, c_or_f)
    s = orc.initsolver(cbc, np.asarray(bc), c_or_f)
    rng = np.random.default_rng(5 + cc)
    rhs = IF.halo(ng, rng.uniform(-1, 1, tuple(ng)))
    f = sim.ctx.field()
    f.set(rhs)
    plan.solve_helmholtz(f, alpha)
    sim.ctx.sync()
    ref = rhs.copy(order="F")
    orc.solve_helmholtz(ref, alpha, s)
    got = f.numpy()
    q = 1 if (c_or_f[2] == "f" and cbc[5] == "D") else 0  # the last face row is not solved (solver.f90:77)
    a = got[1:-1, 1:-1, 1:ng[2] + 1 - q]
    b = ref[1:-1, 1:-1, 1:ng[2] + 1 - q]
    err = float(np.abs(a - b).max() / np.abs(b).max())
    print("helmholtz rel err %.3e" % err, flush=True)
    plan.close()
    assert err < 1e-12


# Face-centred non-periodic transforms (fft.f90:80-82, 289-311): walls normal to the staggered
# direction of a velocity component (RODFT00 on the n-1 interior faces for DD, REDFT10/01 for
# ND, RODFT01/10 for DN), with non-zero boundary values and a stretched z.
WALLS = [("fcc", "DDPPDD", (0.3, -0.2, 0.0, 0.0, 0.5, -0.7), (32, 24, 20)),
         ("ccf", "NNDDDD", (0.1, 0.2, 0.4, -0.1, 0.25, -0.3), (32, 24, 20)),
         ("cfc", "PPDDNN", (0.0, 0.0, -0.6, 0.35, 0.2, -0.15), (32, 40, 16)),
         ("fcc", "DNDDDD", (0.2, 0.1, 0.3, -0.3, 0.1, 0.4), (64, 24, 20)),
         ("fcc", "NDNNDD", (0.2, 0.1, 0.3, -0.3, 0.1, 0.4), (48, 16, 12)),
         ("cfc", "DDDNDD", (0.5, -0.5, 0.2, 0.7, 0.0, 0.0), (16, 32, 24)),
         ("cfc", "NNNDNN", (0.5, -0.5, 0.2, 0.7, 0.3, 0.1), (16, 32, 24))]


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
@pytest.mark.parametrize("c_or_f,cbc,bc,ng", WALLS, ids=["%s-%s" % w[:2] for w in WALLS])
def test_helmholtz_face_centred_walls(c_or_f, cbc, bc, ng, precision):
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    from cans_amd import _lib as LB
    from cans_amd.host import Plan, Simulation
    print("\n[case] helmholtz walls %s %s ng=%s %s" % (c_or_f, cbc, ng, precision), flush=True)
    pre = "".join("P" if t == "P" else "N" for t in cbc)
    p = dict(ng=ng, l=(1.1, 0.8, 1.0), gtype=1, gr=1.5, cbcvel=pre * 3, cbcpre=pre, visci=60.0)
    mixed = precision == "mixed"
    sim = Simulation(cases.host_case(p), precision=LB.CANS_MIXED if mixed else LB.CANS_FP64)
    orc = cases.oracle_case(p)
    plan = Plan(sim.ctx, sim.grid, cbc, bc, c_or_f)
    s = orc.initsolver(cbc, np.asarray(bc), c_or_f)
    assert plan.normfft == s["normfft"]
    alpha = -0.013
    rng = np.random.default_rng(8)
    rhs = IF.halo(ng, rng.uniform(-1, 1, tuple(ng)))
    if mixed:  # the FP32 field: both sides start from the same values
        rhs = rhs.astype(np.float32).astype(np.float64)
    f = sim.ctx.field(kind=LB.CANS_FIELD_PRESSURE if mixed else LB.CANS_FIELD_VELOCITY)
    f.set(rhs)
    plan.solve_helmholtz(f, alpha)
    sim.ctx.sync()
    ref = rhs.copy(order="F")
    orc.solve_helmholtz(ref, alpha, s)
    a, b = f.interior(), ref[1:-1, 1:-1, 1:-1]
    q = 1 if (c_or_f[2] == "f" and cbc[5] == "D") else 0  # the last z face is not solved (solver.f90:77)
    a, b = a[:, :, :ng[2] - q], b[:, :, :ng[2] - q]
    err = float(np.abs(a - b).max() / np.abs(b).max())
    print("helmholtz rel err %.3e" % err, flush=True)
    plan.close()
    assert err < (1e-6 if mixed else 1e-12)
