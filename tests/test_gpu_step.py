"""cans_rk3_step (one whole RK3 step, each substep's projection folded into the next
substep's momentum kernel) against three cans_rk3_substep calls, after several steps.

FP64: the same fields, bit for bit.  Mixed mode: compared at 1e-6 of each field's largest value
(the mixed mode's bar against the FP64 oracle; the oracle comparison of the folded mixed step is
in the mixed-mode tests).  With the intermediate velocity as FP32 increments
(cans_ctx_set_mixed_increments) the folded step stores them over the step's initial velocity
(the corrected velocity of substeps 1 and 2 is never stored), the substep calls over the
corrected one, so they round differently.

The cases cover every boundary rule the fold reproduces: periodic directions (the ghost
copies bounduvw makes after correc and the ghosts that keep correc's own value), walls with
Dirichlet and Neumann velocity (tangential ghost rules, the normal component kept), a moving
lid (non-zero wall values), Dirichlet pressure faces, bulk forcing, stretched z, and FP64
refinement of the last substep.  A case with scalars takes the unfolded path.
"""
import numpy as np
import pytest

from cans_amd import initflow as IF
from tests import cases

pytestmark = pytest.mark.gpu

CASES = [("channel", (64, 32, 48)), ("tgv", (32, 32, 32)), ("duct", (32, 16, 24)), ("box", (16, 16, 12)),
         ("open_nd", (16, 24, 12)), ("cavity_dd", (32, 32, 24)), ("ldc", (2, 64, 64)), ("open_x", (64, 16, 12))]


def run_pair(name, ng, precision, nsteps=3, dt=None, refine=None, increments=False):
    import torch
    from cans_amd import _lib as LB
    from cans_amd.host import Simulation
    print("\n[case] rk3_step vs substeps %s %s %s" % (name, ng, precision), flush=True)
    p = cases.params(name, ng=ng)
    prec = LB.CANS_MIXED if precision == "mixed" else LB.CANS_FP64
    sims = [Simulation(cases.host_case(p), precision=prec) for _ in range(2)]
    init = cases.initial_state(name, p, sims[0].grid, seed=5)
    if prec == LB.CANS_MIXED:
        init = [q.astype(np.float32).astype(np.float64) for q in init]
    sc = cases.initial_scalars(name, p, sims[0].grid, seed=6) if p.get("scalars") else []
    for sim in sims:
        sim.set_state(*(IF.halo(ng, q) for q in init), s=[IF.halo(ng, q) for q in sc])
        sim.init_bounds()
        if refine:
            sim.set_refinement(*refine)
        if increments:
            sim.set_mixed_increments(True)
    dt = dt or (2e-4 if name == "channel" else 1e-3)
    for _ in range(nsteps):
        sims[0].rk3_step(dt)
        sims[1].rk3_step_substeps(dt)
    torch.cuda.synchronize()
    exact = prec == LB.CANS_FP64
    # pp, the last substep's correction, is compared bit for bit in FP64; in the mixed mode it is
    # a small difference of nearly equal fields whose FP32 noise differs between the two paths
    for k in ("u", "v", "w", "p", "pp") if exact else ("u", "v", "w", "p"):
        a, b = getattr(sims[0], k).numpy(), getattr(sims[1], k).numpy()
        assert np.isfinite(a).all(), k
        if exact:
            np.testing.assert_array_equal(a, b, err_msg=k)
        else:
            a, b = a[1:-1, 1:-1, 1:-1], b[1:-1, 1:-1, 1:-1]
            if k == "p" and "D" not in p["cbcpre"]:  # singular: compare up to the constant
                wgt = np.broadcast_to(1 / sims[0].grid.dzfi[1:-1], a.shape)
                a = a - (a * wgt).sum() / wgt.sum()
                b = b - (b * wgt).sum() / wgt.sum()
            err = np.abs(a - b).max() / max(np.abs(b).max(), 1e-300)
            print("%s rel diff %.3e" % (k, err), flush=True)
            assert err <= 1e-6, (k, err)
    for i, (a, b) in enumerate(zip(sims[0].rhs_o, sims[1].rhs_o)):
        a, b = a.cpu().numpy(), b.cpu().numpy()
        if exact:
            np.testing.assert_array_equal(a, b, err_msg="rhs_o %d" % i)
        else:
            assert np.abs(a - b).max() <= 1e-5 * max(np.abs(b).max(), 1e-300), i
    for i in range(len(sc)):
        np.testing.assert_array_equal(sims[0].s[i].numpy(), sims[1].s[i].numpy(), err_msg="s%d" % i)
    for sim in sims:
        sim.close()


@pytest.fixture(scope="module")
def torch():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    return torch


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
@pytest.mark.parametrize("name,ng", CASES, ids=["%s-%dx%dx%d" % ((c,) + g) for c, g in CASES])
def test_rk3_step_equals_substeps(torch, name, ng, precision):
    run_pair(name, ng, precision)


def test_rk3_step_refined_last_substep(torch):
    run_pair("channel", (64, 32, 48), "mixed", refine=(2, 4))


@pytest.mark.parametrize("name,ng", [("channel", (64, 32, 48)), ("cavity_dd", (32, 32, 24)), ("ldc", (2, 64, 64))])
def test_rk3_step_mixed_increments(torch, name, ng):
    """The intermediate velocity as FP32 increments over the step's velocity: the folded substeps
    store theirs over the step's initial velocity, the unfolded ones over the corrected one."""
    run_pair(name, ng, "mixed", increments=True)


def test_rk3_step_with_scalars_unfolded(torch):
    run_pair("rb_channel", (32, 16, 24), "fp64", dt=1e-3)
