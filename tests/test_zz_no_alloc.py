"""cans_solver_noalloc (the allocation-free Poisson solve on the library's fixed device
scratch) against cans_solver, the entry point it mirrors, through the C ABI.

Two calls in a row on two contexts, the second grid smaller than the first, so the
second call reuses the scratch the first one left behind: both must equal cans_solver's
result bit for bit (the same kernels on the same inputs), in FP64 and in the mixed
mode.  A grid whose solve needs more than the fixed scratch is refused.
"""
import ctypes as C

import numpy as np
import pytest

from cans_amd import initflow as IF
from tests import cases

pytestmark = pytest.mark.gpu

# (case, grid): a 128^3 tri-periodic box (the largest solve the scratch covers in FP64,
# periodic z), then a smaller stretched channel (NN walls in z)
CALLS = [("tgv", (128, 128, 128)), ("channel", (32, 16, 24))]


def _solve(sim, fn, rhs):
    from cans_amd import _lib as LB
    pl = sim.plan
    f = sim.pp
    f.set(rhs)
    n = LB.ivec(sim.ctx.n)
    rc = getattr(LB.lib(), fn)(sim.ctx.h, n, n, pl.h, pl.normfft, C.c_void_p(pl.lam.data_ptr()),
                               C.c_void_p(pl.a.data_ptr()), C.c_void_p(pl.b.data_ptr()),
                               C.c_void_p(pl.c.data_ptr()), pl.cbc.encode(), pl.c_or_f.encode(), C.c_void_p(f.ptr))
    LB.check(rc, fn)
    sim.ctx.sync()
    return f.numpy().copy()


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_noalloc_solver_matches_cans_solver(precision):
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    from cans_amd import _lib as LB
    from cans_amd.host import Simulation
    prec = LB.CANS_FP64 if precision == "fp64" else LB.CANS_MIXED
    for name, ng in CALLS:
        print("\n[case] noalloc solve %s %s %s" % (precision, name, ng), flush=True)
        p = cases.params(name, ng=ng)
        sim = Simulation(cases.host_case(p), precision=prec)
        rng = np.random.default_rng(5)
        rhs = rng.uniform(-1, 1, tuple(ng))
        wgt = np.broadcast_to(1 / sim.grid.dzfi[1:-1], rhs.shape)
        rhs -= (rhs * wgt).sum() / wgt.sum()
        h = IF.halo(ng, rhs)
        a = _solve(sim, "cans_solver_noalloc", h)
        b = _solve(sim, "cans_solver", h)
        assert np.isfinite(a).all()
        np.testing.assert_array_equal(a, b)
        sim.close()


def test_noalloc_solver_refuses_grids_above_the_scratch():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    from cans_amd import _lib as LB
    from cans_amd.host import Simulation
    ng = (256, 128, 128)  # FP64 tri-periodic: about 78 MB of scratch, above 64 MiB
    p = cases.params("tgv", ng=ng)
    sim = Simulation(cases.host_case(p))
    pl = sim.plan
    n = LB.ivec(ng)
    rc = LB.lib().cans_solver_noalloc(sim.ctx.h, n, n, pl.h, pl.normfft, C.c_void_p(pl.lam.data_ptr()),
                                      C.c_void_p(pl.a.data_ptr()), C.c_void_p(pl.b.data_ptr()),
                                      C.c_void_p(pl.c.data_ptr()), pl.cbc.encode(), pl.c_or_f.encode(),
                                      C.c_void_p(sim.pp.ptr))
    assert rc == LB.CANS_EINVAL
    assert b"scratch" in LB.lib().cans_last_error()
    sim.close()
