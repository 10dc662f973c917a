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
