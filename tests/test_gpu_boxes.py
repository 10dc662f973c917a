"""Independent Poisson boxes of different sizes solved together (cans_poisson_batch_*),
box by box against the oracle's solver (solver.f90:17-112 on each box).

* lengths 2^m, 3*2^m, and (second set) 5*2^m, 9*2^m, 15*2^m (radix-3 and radix-5 stages)
  along x and y, PP/NN/DD pairs, stretched z; a third set with ND/DN pairs along x and y;
* FP64: every box at 1e-12 relative L-inf against the oracle, and bit for bit against
  the single-box cans_solver of the same box (the same per-line arithmetic in other
  launch shapes);
* the mixed mode (FP32 storage) at 1e-6 after gauge fixing;
* the 64-box workload (nx, ny from {32, 48, 64, 96, 128}, nz in [32, 160], stretching
  in [0, 2]) in FP64 against the oracle;
* a HIP graph of the batch solve replayed three times, equal to the eager solve.

Boxes with no Dirichlet wall are singular: the zero-wavenumber line is pinned only
when its last pivot is within one ulp of zero (solver.f90:151-160), which on most
stretched grids with many z cells it is not, and the solution then carries a
constant set by rounding (the oracle with another FFT factorisation differs from
itself by up to 2e-3 before and 1e-15 after removing it).  Those boxes are compared
with their volume-weighted mean removed.
"""
import ctypes as C

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


def gauge(a, box, dzf):
    """a with its volume-weighted mean removed when the box has no Dirichlet wall."""
    if "D" in box["cbc"]:
        return a
    wgt = np.broadcast_to(dzf, a.shape)
    return a - (a * wgt).sum() / wgt.sum()


def oracle_box(box):
    from oracle import oracle as O
    return O.OracleCase(ng=box["n"], l=box["l"], gtype=box["gtype"], gr=box["gr"], cbcvel="PP" * 9,
                        cbcpre=box["cbc"], visci=100.0)


MIX_BOXES = cases.poisson_boxes(10, seed=5, sizes=(24, 32, 48, 64, 96), nz=(8, 40),
                                cbcs=("PPPPNN", "NNPPDN", "PPDDNN", "DDNNNN", "PPPPDD"))
ODD_BOXES = cases.poisson_boxes(8, seed=6, sizes=(40, 72, 80, 120, 160), nz=(8, 32),
                                cbcs=("PPPPNN", "NNPPDN", "PPDDNN", "DDNNNN", "PPPPDD"))
# REDFT11 / RODFT11 along x and y (pressure ND, DN)
NDDN_BOXES = cases.poisson_boxes(6, seed=13, sizes=(16, 24, 32, 40, 64), nz=(8, 24),
                                 cbcs=("NDPPNN", "PPDNDD", "DNNDNN", "NDNDDD"))
BOX_SETS = {"mix": MIX_BOXES, "odd": ODD_BOXES, "nd": NDDN_BOXES}


def run_batch(boxes, precision, seed):
    from cans_amd import _lib as LB
    from cans_amd.host import Context, PoissonBatch
    big = tuple(max(b["n"][d] for b in boxes) for d in range(3))
    ctx = Context(big, precision=precision)
    batch = PoissonBatch(ctx, boxes)
    rng = np.random.default_rng(seed)
    rhs = []
    for i, b in enumerate(boxes):
        r = cases.box_rhs(b, 1 / batch.dzfi[i][1:-1], rng)
        if precision == LB.CANS_MIXED:
            r = r.astype(np.float32).astype(np.float64)
        rhs.append(IF.halo(b["n"], r))
        batch.set(i, rhs[-1])
    batch.solve()
    ctx.sync()
    return ctx, batch, rhs


@pytest.mark.parametrize("which", list(BOX_SETS))
def test_boxes_fp64_match_oracle_and_single_solver(torch, which):
    from cans_amd import _lib as LB
    from cans_amd.host import Simulation
    boxes = BOX_SETS[which]
    print("\n[case] %d boxes fp64: %s" % (len(boxes), [(b["n"], b["cbc"]) for b in boxes]), flush=True)
    ctx, batch, rhs = run_batch(boxes, LB.CANS_FP64, seed=3)
    for i, b in enumerate(boxes):
        orc = oracle_box(b)
        ref = rhs[i].copy(order="F")
        orc.solver(ref)
        assert batch.normfft[i] == orc.p_solver["normfft"], (i, batch.normfft[i], orc.p_solver["normfft"])
        got = batch.get(i)
        dzf = 1 / batch.dzfi[i][1:-1]
        err = relerr(gauge(got[1:-1, 1:-1, 1:-1], b, dzf), gauge(ref[1:-1, 1:-1, 1:-1], b, dzf))
        print("box %d %s %s rel err %.3e" % (i, b["n"], b["cbc"], err), flush=True)
        assert err < RTOL, (i, b, err)
        if i < 3:  # the single-box solver on the same box (radix-8 stages, as the batch): identical bits
            p = dict(ng=b["n"], l=b["l"], gtype=1, gr=b["gr"], cbcvel="PP" * 9, cbcpre=b["cbc"], visci=100.0)
            sim = Simulation(cases.host_case(p))
            assert LB.lib().cans_fft_set_radix(sim.plan.h, (C.c_int * 4)(8, 8, 8, 8)) == LB.CANS_OK
            sim.pp.set(rhs[i])
            sim.solver()
            sim.ctx.sync()
            np.testing.assert_array_equal(sim.pp.numpy()[1:-1, 1:-1, 1:-1], got[1:-1, 1:-1, 1:-1],
                                          err_msg="box %d vs cans_solver" % i)
            sim.close()
    batch.close()


@pytest.mark.parametrize("which", list(BOX_SETS))
def test_boxes_mixed(torch, which):
    from cans_amd import _lib as LB
    boxes = BOX_SETS[which]
    print("\n[case] %d boxes mixed" % len(boxes), flush=True)
    ctx, batch, rhs = run_batch(boxes, LB.CANS_MIXED, seed=4)
    for i, b in enumerate(boxes):
        orc = oracle_box(b)
        ref = rhs[i].copy(order="F")
        orc.solver(ref)
        dzf = 1 / batch.dzfi[i][1:-1]
        err = relerr(gauge(batch.get(i)[1:-1, 1:-1, 1:-1], b, dzf), gauge(ref[1:-1, 1:-1, 1:-1], b, dzf))
        print("box %d %s %s mixed rel err (gauge-fixed) %.3e" % (i, b["n"], b["cbc"], err), flush=True)
        assert err < MIXED_TOL, (i, b, err)
    batch.close()


def test_boxes_64_workload(torch):
    from cans_amd import _lib as LB
    boxes = cases.poisson_boxes(64, seed=2024)
    print("\n[case] 64 boxes, %d cells" % sum(int(np.prod(b["n"])) for b in boxes), flush=True)
    ctx, batch, rhs = run_batch(boxes, LB.CANS_FP64, seed=6)
    worst = 0.0
    for i, b in enumerate(boxes):
        orc = oracle_box(b)
        ref = rhs[i].copy(order="F")
        orc.solver(ref)
        dzf = 1 / batch.dzfi[i][1:-1]
        err = relerr(gauge(batch.get(i)[1:-1, 1:-1, 1:-1], b, dzf), gauge(ref[1:-1, 1:-1, 1:-1], b, dzf))
        worst = max(worst, err)
        assert err < RTOL, (i, b, err)
    print("64 boxes: worst rel err %.3e, x/y work-groups of %d lanes" % (worst, batch.threads()), flush=True)
    assert batch.threads() in (128, 256, 512)
    batch.close()


def test_boxes_graph_replay(torch):
    from cans_amd import _lib as LB
    from cans_amd.host import Context, PoissonBatch
    boxes = MIX_BOXES[:6]
    print("\n[case] graph of %d boxes" % len(boxes), flush=True)
    s = torch.cuda.Stream()
    big = tuple(max(b["n"][d] for b in boxes) for d in range(3))
    ctx = Context(big, precision=LB.CANS_FP64, stream=s)
    batch = PoissonBatch(ctx, boxes)
    rng = np.random.default_rng(8)
    rhs = [IF.halo(b["n"], cases.box_rhs(b, 1 / batch.dzfi[i][1:-1], rng)) for i, b in enumerate(boxes)]

    def load():
        for i in range(len(boxes)):
            batch.set(i, rhs[i])
        torch.cuda.synchronize()

    load()
    with torch.cuda.stream(s):
        batch.solve()
    s.synchronize()
    eager = [batch.get(i) for i in range(len(boxes))]
    g = torch.cuda.CUDAGraph()
    load()
    with torch.cuda.graph(g, stream=s):
        batch.solve()
    for r in range(3):
        load()
        g.replay()
        torch.cuda.synchronize()
        junk = torch.ones(1 << 20, device="cuda")  # eager work between replays
        junk.mul_(2.0).sum().item()
        for i in range(len(boxes)):
            assert np.array_equal(batch.get(i), eager[i]), (r, i)
        print("replay %d identical to eager" % r, flush=True)
    batch.close()
