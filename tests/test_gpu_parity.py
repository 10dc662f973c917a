"""Parity of the HIP path (through the C ABI) against the oracle on identical inputs.

Stencil, boundary and reduction kernels keep the reference's association and
rounding, so they are compared bit for bit.  The Poisson solve differs from the
oracle only through the transforms' rounding; it and whole RK3 steps are held to
the FP64 bar of 1e-12 relative L-inf (max |gpu - oracle| / max |oracle|, taken
over the full halo'd arrays).
"""
import sys

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


def pair(name, ng, seed=7):
    from cans_amd.host import Simulation
    print("\n[case] %s ng=%s seed=%d" % (name, ng, seed), flush=True)
    p = cases.params(name, ng=ng)
    sim = Simulation(cases.host_case(p))
    orc = cases.oracle_case(p)
    u, v, w, pr = cases.initial_state(name, p, sim.grid, seed=seed)
    st = orc.new_state()
    for k, q in zip("uvwp", (u, v, w, pr)):
        st[k][1:-1, 1:-1, 1:-1] = q
    sim.set_state(*(IF.halo(p["ng"], q) for q in (u, v, w, pr)))
    return p, sim, orc, st


SOLVER_CASES = [("tgv", (16, 16, 16)), ("tgv", (32, 8, 12)), ("channel", (32, 16, 24)),
                ("ldc", (2, 64, 64)), ("duct", (16, 32, 20)), ("cavity_dd", (16, 32, 24)),
                ("channel", (128, 64, 48)),
                # radix-3 transform lengths (3 * 2^m) along x and y
                ("channel", (48, 96, 40)), ("cavity_dd", (96, 24, 20)), ("box", (24, 48, 16)),
                # radix-5 and mixed odd factors (5 * 2^m, 15 * 2^m, 9 * 2^m, 25 * 2^m)
                ("channel", (40, 120, 24)), ("cavity_dd", (72, 80, 16)), ("box", (200, 40, 12)),
                # REDFT11 along x and RODFT11 along y (pressure ND, DN)
                ("open_nd", (16, 24, 12)), ("open_nd", (64, 32, 20)), ("open_nd", (40, 12, 10)),
                # DST-II/III along x on lines of 64 elements and more (more than one tile chunk of the
                # work buffer)
                ("open_x", (64, 16, 12)), ("open_x", (96, 24, 10)), ("open_x", (256, 8, 8))]


@pytest.mark.parametrize("name,ng", SOLVER_CASES, ids=["%s-%dx%dx%d" % ((c,) + g) for c, g in SOLVER_CASES])
def test_solver_matches_oracle(torch, name, ng):
    p, sim, orc, st = pair(name, ng)
    rng = np.random.default_rng(11)
    rhs = rng.uniform(-1, 1, tuple(ng))
    if p["cbcpre"][4:6] != "DD":
        wgt = np.broadcast_to(1 / sim.grid.dzfi[1:-1], rhs.shape)
        rhs -= (rhs * wgt).sum() / wgt.sum()
    f = IF.halo(ng, rhs)
    sim.pp.set(f)
    sim.solver()
    sim.ctx.sync()
    st["pp"][...] = f
    orc.solver(st["pp"])
    got = sim.pp.numpy()[1:-1, 1:-1, 1:-1]
    ref = st["pp"][1:-1, 1:-1, 1:-1]
    raw = relerr(got, ref)
    # with no Dirichlet wall the zero-wavenumber line is pinned only when its last pivot
    # is within one ulp of zero (solver.f90:151-160); otherwise the solution carries a
    # constant set by the transforms' rounding, which the comparison removes
    if "D" not in p["cbcpre"]:
        wgt = np.broadcast_to(1 / sim.grid.dzfi[1:-1], got.shape)
        got = got - (got * wgt).sum() / wgt.sum()
        ref = ref - (ref * wgt).sum() / wgt.sum()
    err = relerr(got, ref)
    print("solver rel err %.3e (before removing the mean %.3e)" % (err, raw), flush=True)
    assert err < RTOL


BOUND_CASES = [("tgv", (8, 4, 5)), ("channel", (8, 4, 5)), ("ldc", (2, 8, 6)), ("cavity_dd", (4, 8, 5))]
FLAGS = [(-1, -99, -99), (0, -1, -99), (1, 1, 0), (0, -99, -99), (1, -99, -99)]


@pytest.mark.parametrize("name,ng", BOUND_CASES, ids=[c for c, g in BOUND_CASES])
def test_bounds_bitwise(torch, name, ng):
    p, sim, orc, st = pair(name, ng)
    rng = np.random.default_rng(3)
    bcvel = rng.uniform(-1, 1, 18)  # non-zero wall values exercise the BC factors
    sim.case.bcvel = tuple(bcvel)
    orc.bcvel = bcvel.copy()
    for flags in FLAGS:
        fields = [np.asfortranarray(rng.uniform(-1, 1, tuple(x + 2 for x in ng))) for _ in range(4)]
        for f, a in zip((sim.u, sim.v, sim.w, sim.p), fields):
            f.set(a)
        ou, ov, ow, op = (a.copy(order="F") for a in fields)
        sim.bounduvw(sim.u, sim.v, sim.w, *flags)
        orc.bounduvw(ou, ov, ow, *flags)
        for hs in (-99, 1, -1):
            sim.boundp(sim.p, hs)
            orc.boundp(op, hs)
        sim.ctx.sync()
        for f, ref in zip((sim.u, sim.v, sim.w, sim.p), (ou, ov, ow, op)):
            np.testing.assert_array_equal(f.numpy(), ref, err_msg="flags=%s" % (flags,))


@pytest.mark.parametrize("name,ng", [("tgv", (16, 8, 10)), ("channel", (32, 16, 20))])
def test_operators_bitwise(torch, name, ng):
    """rk (mom_xyz_ad + update + forcing), fillps, correc, updatep, chkdiv, chkdt."""
    p, sim, orc, st = pair(name, ng)
    rng = np.random.default_rng(5)
    for k in "uvwp":
        a = np.asfortranarray(rng.uniform(-1, 1, tuple(x + 2 for x in ng)))
        st[k][...] = a
        getattr(sim, k).set(a)
    ro = [np.asfortranarray(rng.uniform(-1, 1, tuple(ng))) for _ in range(3)]
    for t, a in zip(sim.rhs_o, ro):
        t.copy_(torch.from_numpy(a.ravel(order="F")))
    st["dudtrko"][...], st["dvdtrko"][...], st["dwdtrko"][...] = ro
    # the first rk call zeroes RHS_old in both (rk.f90:72-80)
    for name_ in ("dudtrko", "dvdtrko", "dwdtrko"):
        st[name_][...] = 0.0
    dt = 1e-3
    for irk in (1, 2):
        f_gpu = sim.rk(irk, dt)
        f_orc = orc.rk(cases_rk(irk), dt, st["p"], st["dudtrko"], st["dvdtrko"], st["dwdtrko"],
                       st["u"], st["v"], st["w"])
        sim.ctx.sync()
        for c, k in enumerate("uvw"):
            if p.get("is_forced", (0, 0, 0))[c]:
                # f = velf - mean: the mean is a reduction whose summation order differs (1 ulp in f)
                np.testing.assert_allclose(getattr(sim, k).numpy(), st[k], rtol=0, atol=1e-15 * np.abs(st[k]).max())
            else:
                np.testing.assert_array_equal(getattr(sim, k).numpy(), st[k], err_msg="rk irk=%d %s" % (irk, k))
        for t, name_ in zip(sim.rhs_o, ("dudtrko", "dvdtrko", "dwdtrko")):
            np.testing.assert_array_equal(t.cpu().numpy(), st[name_].ravel(order="F"))
        np.testing.assert_allclose(f_gpu, f_orc, rtol=1e-13, atol=1e-15)
    dti = 1.0 / (dt * 8 / 60)
    sim.fillps(dti)
    orc.fillps(st["u"], st["v"], st["w"], st["pp"], dti)
    sim.ctx.sync()
    if any(p.get("is_forced", (0, 0, 0))):  # (u(i) - u(i-1)) amplifies the 1-ulp forcing difference
        assert relerr(sim.pp.numpy()[1:-1, 1:-1, 1:-1], st["pp"][1:-1, 1:-1, 1:-1]) < 1e-12
    else:
        np.testing.assert_array_equal(sim.pp.numpy()[1:-1, 1:-1, 1:-1], st["pp"][1:-1, 1:-1, 1:-1])
    pp = np.asfortranarray(rng.uniform(-1, 1, tuple(x + 2 for x in ng)))
    sim.pp.set(pp)
    st["pp"][...] = pp
    sim.correc(dt)
    orc.correc(st["pp"], st["u"], st["v"], st["w"], dt)
    sim.updatep()
    orc.updatep(st["pp"], st["p"])
    sim.ctx.sync()
    for c, k in enumerate("uvwp"):
        if c < 3 and p.get("is_forced", (0, 0, 0))[c]:
            np.testing.assert_allclose(getattr(sim, k).numpy(), st[k], rtol=0, atol=1e-15 * np.abs(st[k]).max())
        else:
            np.testing.assert_array_equal(getattr(sim, k).numpy(), st[k], err_msg="correc/updatep " + k)
    tot, mx = sim.chkdiv()
    otot, omx = orc.chkdiv(st["u"], st["v"], st["w"])
    assert mx == omx
    assert abs(tot - otot) <= 1e-13 * abs(otot)
    dtm = sim.chkdt()
    odtm = orc.chkdt(st["u"], st["v"], st["w"])
    assert dtm == odtm


def pressure_err(p, sim, st):
    """Relative L-inf error of the pressure interior, up to the additive constant when the
    pressure operator is singular (no Dirichlet pressure boundary).  The reference fixes that
    constant with a 1-ulp pivot test (solver.f90:156-160) which, on most stretched or periodic
    z grids, does not fire: the constant is then round-off of the zero-wavenumber rhs divided by
    a near-zero pivot, and differs between any two transform implementations."""
    a = sim.p.numpy()[1:-1, 1:-1, 1:-1]
    b = st["p"][1:-1, 1:-1, 1:-1]
    if "D" in p["cbcpre"]:
        return relerr(a, b), 0.0
    wgt = np.broadcast_to(1 / sim.grid.dzfi[1:-1], a.shape)
    ma = (a * wgt).sum() / wgt.sum()
    mb = (b * wgt).sum() / wgt.sum()
    return relerr(a - ma, b - mb), ma - mb


def cases_rk(irk):
    from oracle.oracle import RKCOEFF
    return RKCOEFF[irk - 1]


# the walled box (no periodic x) takes the unfused divergence path of the fused substep
STEP_CASES = [("tgv", (32, 32, 32), 1e-3, 3), ("channel", (64, 32, 48), 2e-4, 3), ("duct", (32, 16, 24), 1e-3, 2),
              ("cavity_dd", (16, 16, 12), 1e-3, 2), ("box", (16, 16, 12), 1e-3, 2),
              ("channel", (96, 48, 32), 2e-4, 2), ("channel", (80, 120, 24), 2e-4, 2),
              ("open_nd", (16, 24, 12), 1e-3, 2)]


@pytest.mark.parametrize("name,ng,dt,nsteps", STEP_CASES, ids=["%s-%dx%dx%d" % ((c[0],) + c[1]) for c in STEP_CASES])
@pytest.mark.parametrize("fused", [True, False], ids=["fused", "per-routine"])
def test_rk3_steps_match_oracle(torch, name, ng, dt, nsteps, fused):
    """Velocity at 1e-12; pressure (gauge-fixed) at 1e-12 or, where the stretched
    z-operator's near-zero wavenumbers make that unattainable for any two FP64
    transform implementations, within the oracle's own spread: the same oracle run
    with an alternative FFT factorisation (orc_set_fft_variant(1))."""
    from oracle import oracle as O
    p, sim, orc, st = pair(name, ng)
    st2 = {k: v.copy(order="F") for k, v in st.items()}
    sim.init_bounds()
    orc.init_bounds(st)
    orc.init_bounds(st2)
    for it in range(nsteps):
        (sim.rk3_step if fused else sim.rk3_step_unfused)(dt)
        orc.rk3_step(st, dt)
    O.fft_variant(1)
    try:
        for it in range(nsteps):
            orc.rk3_step(st2, dt)
    finally:
        O.fft_variant(0)
    sim.ctx.sync()
    for k in "uvw":
        err = relerr(getattr(sim, k).numpy(), st[k])
        spread = relerr(st2[k], st[k])
        print("%s rel err %.3e (oracle spread %.3e)" % (k, err, spread), flush=True)
        assert err < RTOL, (k, err)
    err, offset = pressure_err(p, sim, st)

    class _S:
        pass
    s2 = _S()
    s2.p = _S()
    s2.p.numpy = lambda: st2["p"]
    s2.grid = sim.grid
    spread, _ = pressure_err(p, s2, st)
    print("p rel err (gauge-fixed) %.3e, oracle spread %.3e, constant offset %.3e" % (err, spread, offset),
          flush=True)
    assert err < max(RTOL, 4 * spread), ("p", err, spread)
    tot, mx = sim.chkdiv()
    otot, omx = orc.chkdiv(st["u"], st["v"], st["w"])
    print("divmax gpu %.3e oracle %.3e" % (mx, omx), flush=True)
    assert mx < 2.2e-9 and mx <= 10 * max(omx, 1e-15)


FUSED_CASES = [("channel", (64, 32, 48), 2e-4), ("tgv", (32, 32, 32), 1e-3), ("box", (16, 16, 12), 1e-3),
               ("open_nd", (16, 24, 12), 1e-3)]


@pytest.mark.parametrize("name,ng,dt", FUSED_CASES, ids=["%s-%dx%dx%d" % ((c[0],) + c[1]) for c in FUSED_CASES])
def test_fused_substep_bitwise_per_routine(torch, name, ng, dt):
    """FP64: cans_rk3_substep (divergence fused into the first transform pass where x is
    periodic, forcing applied where the velocity is read, updatep fused into correc) gives
    the same bits as the per-routine sequence of main.f90:437-473 over two RK3 steps."""
    from cans_amd.host import Simulation
    p, sim, orc, st = pair(name, ng)
    sim2 = Simulation(cases.host_case(p))
    u, v, w, pr = cases.initial_state(name, p, sim.grid, seed=7)
    sim2.set_state(*(IF.halo(p["ng"], q) for q in (u, v, w, pr)))
    sim.init_bounds()
    sim2.init_bounds()
    for it in range(2):
        sim.rk3_step(dt)
        sim2.rk3_step_unfused(dt)
    sim.ctx.sync()
    sim2.ctx.sync()
    for k in ("u", "v", "w", "p"):
        np.testing.assert_array_equal(getattr(sim, k).numpy(), getattr(sim2, k).numpy(), err_msg=k)
    print("fused and per-routine steps identical", flush=True)


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_natural_grid_channel_steps(torch, precision):
    """The channel on the natural wall-resolved grid (is_gridpoint_natural_channel, initgrid.f90:46-58,
    162-203): three RK3 steps against the oracle on the same grid."""
    from cans_amd import _lib as LB
    from cans_amd.host import Simulation
    print("\n[case] natural-grid channel %s" % precision, flush=True)
    p = cases.params("channel", ng=(32, 16, 48))
    p["natural_grid"] = True
    mixed = precision == "mixed"
    sim = Simulation(cases.host_case(p), precision=LB.CANS_MIXED if mixed else LB.CANS_FP64)
    orc = cases.oracle_case(p)
    np.testing.assert_allclose(sim.grid.dzf, orc.dzf, rtol=1e-15, atol=0)
    init = cases.initial_state("channel", p, sim.grid, seed=9)
    if mixed:
        init = [q.astype(np.float32).astype(np.float64) for q in init]
    st = orc.new_state()
    for k, q in zip("uvwp", init):
        st[k][1:-1, 1:-1, 1:-1] = q
    sim.set_state(*(IF.halo(p["ng"], q) for q in init))
    sim.init_bounds()
    orc.init_bounds(st)
    for _ in range(3):
        sim.rk3_step(2e-4)
        orc.rk3_step(st, 2e-4)
    sim.ctx.sync()
    for k in "uvw":
        err = relerr(getattr(sim, k).numpy()[1:-1, 1:-1, 1:-1], st[k][1:-1, 1:-1, 1:-1])
        print("%s rel err %.3e" % (k, err), flush=True)
        assert err < (1e-6 if mixed else RTOL), (k, err)


def test_transform_lengths_and_pairs_refused(torch):
    """cans_fftini takes any x/y length up to 8192: 2^m, or 2^m (m >= 3) times at most three
    factors 3 and 5 (ND/DN pairs: twice the length), through the FFT passes, the others as a
    dense matrix; longer non-FFT lengths are refused with CANS_EUNSUPPORTED, boundary pairs
    other than PP, NN, DD, ND, DN with CANS_EINVAL.  The checks run before any device work."""
    import ctypes as C
    from cans_amd import _lib as LB
    L = LB.lib()

    def fftini(ng, bc):
        h = C.c_void_p()
        rc = L.cans_ctx_create(LB.ivec(ng), 0, LB.CANS_FP64, None, C.byref(h))
        assert rc == LB.CANS_OK
        plan, nf = C.c_void_p(), C.c_double()
        rc = L.cans_fftini(h, LB.ivec(ng), bc, b"ccc", C.byref(plan), C.byref(nf))
        msg = L.cans_last_error()
        if rc == LB.CANS_OK:
            L.cans_fftend(plan)
        L.cans_ctx_destroy(h)
        return rc, msg

    for ng, bc, want in [((14, 16, 8), b"PPPPNN", LB.CANS_OK),             # 7 * 2: dense x
                         ((12, 16, 8), b"PPPPNN", LB.CANS_OK),             # 3 * 4: m < 3
                         ((16, 270, 8), b"PPPPNN", LB.CANS_OK),            # 27 * 10: four odd factors
                         ((8200, 2, 8), b"PPPPNN", LB.CANS_EUNSUPPORTED),  # 8 * 1025: dense and too long
                         ((16, 8, 8), b"PNPPNN", LB.CANS_EINVAL),
                         ((12, 16, 8), b"NDPPNN", LB.CANS_OK),             # REDFT11 on a line of 24
                         ((16, 120, 8), b"PPDNNN", LB.CANS_OK),            # RODFT11 on a line of 240
                         ((40, 72, 8), b"NNDDNN", LB.CANS_OK)]:
        rc, msg = fftini(ng, bc)
        print("[case] fftini %s %s -> %d %s" % (ng, bc.decode(), rc, msg.decode() if rc else ""), flush=True)
        assert rc == want, (ng, bc, rc, msg)
