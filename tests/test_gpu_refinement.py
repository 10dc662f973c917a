"""Poisson solves with FP64 iterative refinement (cans_solver_refined) against the oracle.

In a CANS_MIXED context the inner solves use FP32 transforms and storage; the
residual b - L x and the solution stay FP64.  On stretched channels (singular
Neumann z), a duct and a Dirichlet cavity, the relative error against the oracle's
FP64 solve (gauge-fixed when singular) must fall from the FP32 level (~1e-7)
without refinement to the FP64 level after two sweeps (on the stretched channels
the refined FP64 solution and the oracle's differ by the operator's conditioning
times the FP64 rounding, ~5e-13; FP64 and mixed contexts converge to the same
digits); in an FP64 context refinement stays at that level.
"""
import numpy as np
import pytest

from cans_amd import initflow as IF
from tests import cases

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def torch():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    return torch


def gauge(a, dzf, singular):
    if not singular:
        return a
    wgt = np.broadcast_to(dzf, a.shape)
    return a - (a * wgt).sum() / wgt.sum()


CASES = [("channel", (64, 32, 48)), ("channel", (128, 64, 96)), ("duct", (16, 32, 20)), ("cavity_dd", (16, 32, 24))]


@pytest.mark.parametrize("name,ng", CASES, ids=["%s-%dx%dx%d" % ((c,) + g) for c, g in CASES])
def test_refinement_reaches_fp64(torch, name, ng):
    from cans_amd import _lib as LB
    from cans_amd.host import Simulation
    print("\n[case] refined solve %s %s" % (name, ng), flush=True)
    p = cases.params(name, ng=ng)
    orc = cases.oracle_case(p)
    rng = np.random.default_rng(17)
    rhs = rng.uniform(-1, 1, tuple(ng))
    singular = "D" not in p["cbcpre"]
    dzf = 1 / orc.dzfi[1:-1]
    rhs = gauge(rhs, dzf, singular)
    ref = IF.halo(ng, rhs)
    orc.solver(ref)
    ref = gauge(ref[1:-1, 1:-1, 1:-1], dzf, singular)
    errs = {}
    for prec in (LB.CANS_MIXED, LB.CANS_FP64):
        sim = Simulation(cases.host_case(p), precision=prec)
        f = sim.ctx.field()  # FP64 (velocity kind) in both modes
        for sweeps in (0, 1, 2):
            f.set(IF.halo(ng, rhs))
            sim.solver_refined(f, sweeps)
            sim.ctx.sync()
            got = gauge(f.interior(), dzf, singular)
            errs[(prec, sweeps)] = float(np.abs(got - ref).max() / np.abs(ref).max())
        sim.close()
    print("mixed: %s; fp64: %s" % (["%.2e" % errs[(LB.CANS_MIXED, s)] for s in (0, 1, 2)],
                                   ["%.2e" % errs[(LB.CANS_FP64, s)] for s in (0, 1, 2)]), flush=True)
    assert errs[(LB.CANS_MIXED, 0)] < 1e-5
    assert errs[(LB.CANS_MIXED, 2)] < 1e-11
    assert errs[(LB.CANS_MIXED, 2)] < 1e-3 * errs[(LB.CANS_MIXED, 0)]
    for s in (0, 1, 2):
        assert errs[(LB.CANS_FP64, s)] < 1e-11
