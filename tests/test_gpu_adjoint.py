"""Adjoint of the RK3 substep on the GPU (cans_mom_jt, cans_rk3_substep_adjoint,
cans_amd.adjoint) against the oracle's adjoint (oracle/adjoint.py) and central
differences.

* cans_mom_jt: J(q)^T g + base at 1e-13 relative against the oracle's mom_jt, on a
  box with different spacings in x, y and z;
* one adjoint substep (projection, the f1/f2 split, J^T) at 1e-12 against the oracle;
* the gradient of the kinetic energy after 5 steps at 16^3 (checkpoints every 2) at
  1e-10 against the oracle's gradient;
* the Taylor-Green workload: 256^3, nu = 1/1600, dt = 5e-4, 32 steps, checkpoints every
  4, through torch.autograd; the directional derivative along 8 seeded random
  directions against central differences of the GPU's own steps (eps = 1e-6) to 1e-5;
* walls in z on a tanh-stretched grid with bulk forcing (the channel): one adjoint
  substep at 1e-12 against the oracle, with a moving top wall and a Neumann v wall so
  that both ghost folds act; the gradient after 4 steps at 1e-10 against the oracle's;
  the same for walls in y and z (a forced duct) and on all six faces (a closed box with
  a moving lid, DCT pressure solves along x and y), and for 3-D implicit diffusion on a
  forced tri-periodic box (a Helmholtz solve per component in each adjoint substep);
* CANS_MIXED (FP32 pressure side in the forward and adjoint projections): the gradient
  after 4 steps of the Taylor-Green box and of the channel against the FP64 oracle's;
* central differences of the GPU's own steps at larger sizes: a closed 96^3 box with a
  moving lid (8 steps) and a 128^3 tri-periodic box with 3-D implicit diffusion (8 steps),
  3 directions each, to 1e-5;
* scalars (cans_rk3_substep_adjoint_scal): a heated channel with a buoyant scalar
  (Dirichlet walls, a source, gravity along x and z) and a passive one (Neumann walls),
  both bulk-forced:
  one substep against the oracle at 1e-12 (velocity, scalars and both right-hand sides),
  the gradient after 4 steps at 1e-10, and at 128x64x96 x 8 steps central differences
  along 3 directions to 1e-5; a heated cavity with walls in x and z (substep), and the
  reference's differentially heated cavity at 64x2x64 (gradient);
* refusals: a step with scalars through the velocity-only entry point, implicit
  diffusion next to walls and in its z-only form, each an error that leaves the outputs
  untouched.
  and a 256x128x96 channel (Re_b = 5600, 16 steps, checkpoints every 4) whose
  directional derivatives along 4 seeded directions agree with central differences to
  1e-5.
"""
import ctypes as C

import numpy as np
import pytest

from tests import cases

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def torch():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    return torch


def relerr(a, b):
    return max(float(np.abs(x - y).max()) for x, y in zip(a, b)) / max(float(np.abs(y).max()) for y in b)


def tgv_pair(ng, visci=1600.0):
    from cans_amd.host import Simulation
    p = cases.params("tgv", ng=ng)
    p["visci"] = visci
    sim = Simulation(cases.host_case(p))
    return p, sim, cases.oracle_case(p)


def put(sim, arrays):
    """Fresh DeviceFields holding the given interior arrays with periodic ghost cells."""
    from oracle import adjoint as A
    out = []
    for a in arrays:
        f = sim.ctx.field()
        f.set(A.halo(None, a))
        out.append(f)
    return out


def test_mom_jt_matches_oracle(torch):
    from cans_amd import _lib as LB
    from oracle import adjoint as A
    print("\n[case] cans_mom_jt 16x24x10", flush=True)
    p, sim, orc = tgv_pair((16, 24, 10), visci=37.0)
    rng = np.random.default_rng(1)
    q, g, b = ([rng.uniform(-1, 1, tuple(orc.n)) for _ in range(3)] for _ in range(3))
    fq, fg, fb = put(sim, q), put(sim, g), put(sim, b)
    fo = [sim.ctx.field() for _ in range(3)]
    h = [orc.dli[0], orc.dli[1], orc.dzci[1]]
    LB.check(LB.lib().cans_mom_jt(sim.ctx.h, LB.ivec(orc.n), LB.dvec(h), orc.visc,
                                  *[C.c_void_p(f.ptr) for f in fq + fg + fb + fo]), "cans_mom_jt")
    sim.ctx.sync()
    ref = [x + y for x, y in zip(A.mom_jt(orc, q, g), b)]
    err = relerr([f.interior() for f in fo], ref)
    print("mom_jt rel err %.3e" % err, flush=True)
    assert err < 1e-13


def test_substep_adjoint_matches_oracle(torch):
    from oracle import adjoint as A
    from oracle.oracle import RKCOEFF
    print("\n[case] cans_rk3_substep_adjoint 16x24x10", flush=True)
    p, sim, orc = tgv_pair((16, 24, 10), visci=37.0)
    rng = np.random.default_rng(2)
    q, ub, fb = ([rng.uniform(-1, 1, tuple(orc.n)) for _ in range(3)] for _ in range(3))
    fq, fub, ffb = put(sim, q), put(sim, ub), put(sim, fb)
    fubo, ffbo = [sim.ctx.field() for _ in range(3)], [sim.ctx.field() for _ in range(3)]
    pp = sim.ctx.field()
    dt, irk = 3e-3, 2
    sim.rk3_substep_adjoint(irk, dt, [f.ptr for f in fq], fub, ffb, fubo, ffbo, pp)
    sim.ctx.sync()
    f1, f2 = RKCOEFF[irk - 1][0] * dt, RKCOEFF[irk - 1][1] * dt
    wb = A.project(orc, ub)
    g = [f1 * a + c for a, c in zip(wb, fb)]
    ref_ub = [a + c for a, c in zip(wb, A.mom_jt(orc, q, g))]
    ref_fb = [f2 * a for a in wb]
    e1, e2 = relerr([f.interior() for f in fubo], ref_ub), relerr([f.interior() for f in ffbo], ref_fb)
    print("adjoint substep rel err: velocity %.3e, rhs %.3e" % (e1, e2), flush=True)
    assert e1 < 1e-12 and e2 < 1e-12


def test_gradient_matches_oracle(torch):
    from cans_amd.adjoint import kinetic_energy_after_steps
    from oracle import adjoint as A
    print("\n[case] gradient 16^3, 5 steps", flush=True)
    p, sim, orc = tgv_pair((16, 16, 16), visci=100.0)
    u0 = cases.initial_state("tgv", p, sim.grid, seed=5)[:3]
    dt, nsteps = 5e-3, 5
    E_orc, g_orc = A.grad_ke_after_steps(orc, *u0, dt, nsteps, every=2)
    t0 = [torch.tensor(np.ascontiguousarray(a.transpose(2, 1, 0)), device="cuda", requires_grad=True) for a in u0]
    E = kinetic_energy_after_steps(sim, *t0, dt, nsteps, every=2)
    E.backward()
    g_gpu = [t.grad.cpu().numpy().transpose(2, 1, 0) for t in t0]
    eE = abs(E.item() - E_orc) / E_orc
    eg = relerr(g_gpu, g_orc)
    print("E %.15e (oracle %.15e, rel %.2e); gradient rel err %.3e" % (E.item(), E_orc, eE, eg), flush=True)
    assert eE < 1e-13 and eg < 1e-10


@pytest.mark.timeout(600)
def test_taylor_green_256_gradient_central_differences(torch):
    from cans_amd.adjoint import kinetic_energy_after_steps
    print("\n[case] Taylor-Green 256^3, 32 steps, checkpoints every 4", flush=True)
    p, sim, _ = tgv_pair((256, 256, 256), visci=1600.0)
    u0 = cases.initial_state("tgv", p, sim.grid, seed=1)[:3]
    dt, nsteps, every = 5e-4, 32, 4
    t0 = [torch.tensor(np.ascontiguousarray(a.transpose(2, 1, 0)), device="cuda", requires_grad=True) for a in u0]
    E = kinetic_energy_after_steps(sim, *t0, dt, nsteps, every)
    E.backward()
    grad = [t.grad.detach() for t in t0]
    base = [t.detach() for t in t0]
    gen = torch.Generator(device="cuda").manual_seed(2024)
    eps = 1e-6
    worst = 0.0
    with torch.no_grad():
        for r in range(8):
            d = [torch.randn(x.shape, generator=gen, device="cuda", dtype=torch.float64) for x in base]
            ep = kinetic_energy_after_steps(sim, *(x + eps * y for x, y in zip(base, d)), dt, nsteps, every).item()
            em = kinetic_energy_after_steps(sim, *(x - eps * y for x, y in zip(base, d)), dt, nsteps, every).item()
            fd = (ep - em) / (2 * eps)
            ad = sum(float((g * y).sum()) for g, y in zip(grad, d))
            rel = abs(fd - ad) / abs(ad)
            worst = max(worst, rel)
            print("direction %d: central difference %.12e, adjoint %.12e, rel %.2e" % (r, fd, ad, rel), flush=True)
            assert rel < 1e-5, (r, fd, ad)
    print("E = %.15f, worst rel %.2e" % (E.item(), worst), flush=True)


def channel_params(ng, visci=60.0, cbcvel="PPPPDD" * 3, bcvel=None):
    return dict(ng=ng, l=(2.0, 1.2, 1.0), gtype=1, gr=1.8, cbcvel=cbcvel, cbcpre="PPPPNN",
                bcvel=tuple(bcvel) if bcvel is not None else (0.0,) * 18, visci=visci, is_forced=(1, 0, 0),
                velf=(1.0, 0.0, 0.0))


def channel_pair(p):
    from cans_amd.host import Simulation
    return Simulation(cases.host_case(p)), cases.oracle_case(p)


def substep_vs_oracle(p, seed):
    """One adjoint substep on the GPU against the oracle's, random degrees of freedom."""
    from oracle import adjoint as A
    from oracle import oracle as O
    sim, orc = channel_pair(p)
    rng = np.random.default_rng(seed)

    def dofs():
        a = [rng.uniform(-1, 1, tuple(orc.n)) for _ in range(3)]
        A._clear_wall_faces(orc, a)
        return a

    q, ub, fb = dofs(), dofs(), dofs()
    qh = [O.field(orc.n) for _ in range(3)]
    for x, y in zip(qh, q):
        x[1:-1, 1:-1, 1:-1] = y
    orc.bounduvw(*qh)  # the forward's ghost values, wall velocities included

    def put_halo(arrays, halo):
        out = []
        for a in arrays:
            f = sim.ctx.field()
            f.set(a if halo else np.pad(a, 1))
            out.append(f)
        return out

    fq, fub, ffb = put_halo(qh, True), put_halo(ub, False), put_halo(fb, False)
    fubo, ffbo = [sim.ctx.field() for _ in range(3)], [sim.ctx.field() for _ in range(3)]
    pp = sim.ctx.field()
    dt, irk = 2e-3, 2
    sim.rk3_substep_adjoint(irk, dt, [f.ptr for f in fq], fub, ffb, fubo, ffbo, pp)
    sim.ctx.sync()
    ref_ub, ref_fb = A.substep_adjoint(orc, qh, irk - 1, dt, ub, fb)
    got_ub, got_fb = [f.interior() for f in fubo], [f.interior() for f in ffbo]
    for d in A.wall_dirs(orc):  # the normal wall faces are zero
        idx = [slice(None)] * 3
        idx[d] = -1
        assert np.all(got_ub[d][tuple(idx)] == 0.) and np.all(got_fb[d][tuple(idx)] == 0.)
    e1, e2 = relerr(got_ub, ref_ub), relerr(got_fb, ref_fb)
    print("adjoint substep rel err: velocity %.3e, rhs %.3e" % (e1, e2), flush=True)
    assert e1 < 1e-12 and e2 < 1e-12


def test_substep_adjoint_channel_matches_oracle(torch):
    print("\n[case] cans_rk3_substep_adjoint, channel 16x12x20 (stretched, forced, moving wall)", flush=True)
    bc = np.zeros(18)
    bc[5], bc[10] = 1.0, 0.3
    substep_vs_oracle(channel_params((16, 12, 20), visci=37.0, cbcvel="PPPPDD" + "PPPPND" + "PPPPDD", bcvel=bc), 21)


def walled_params(kind, ng, visci=60.0, is_forced=(0, 0, 0)):
    """duct: periodic x, walls in y and z; box: walls on all six faces; a moving lid (u = 1 at z = lz),
    a Neumann u on the y = 0 wall, tanh-stretched z."""
    bc = np.zeros(18)
    bc[5] = 1.0
    if kind == "duct":
        cbcvel, cbcpre = "PPNDDD" + "PPDDDD" + "PPDDDD", "PPNNNN"
    else:
        cbcvel, cbcpre = "DDNDDD" + "DDDDDD" + "DDDDDD", "NNNNNN"
    return dict(ng=ng, l=(1.3, 1.0, 1.0), gtype=1, gr=1.5, cbcvel=cbcvel, cbcpre=cbcpre, bcvel=tuple(bc),
                visci=visci, is_forced=is_forced, velf=(1.0, 0.0, 0.0))


@pytest.mark.parametrize("kind", ["duct", "box"])
def test_substep_adjoint_walls_xyz_matches_oracle(torch, kind):
    print("\n[case] cans_rk3_substep_adjoint, %s 16x12x20 (stretched, moving lid)" % kind, flush=True)
    substep_vs_oracle(walled_params(kind, (16, 12, 20), visci=37.0,
                                    is_forced=(1, 0, 0) if kind == "duct" else (0, 0, 0)), 22)


def gradient_vs_oracle(p, u0, dt=2e-3, nsteps=4):
    from cans_amd.adjoint import kinetic_energy_after_steps
    from oracle import adjoint as A
    import torch
    sim, orc = channel_pair(p)
    A._clear_wall_faces(orc, u0)
    E_orc, g_orc = A.grad_ke_after_steps(orc, *u0, dt, nsteps, every=2)
    t0 = [torch.tensor(np.ascontiguousarray(a.transpose(2, 1, 0)), device="cuda", requires_grad=True) for a in u0]
    E = kinetic_energy_after_steps(sim, *t0, dt, nsteps, every=2)
    E.backward()
    g_gpu = [t.grad.cpu().numpy().transpose(2, 1, 0) for t in t0]
    eE = abs(E.item() - E_orc) / E_orc
    eg = relerr(g_gpu, g_orc)
    print("E %.15e (oracle %.15e, rel %.2e); gradient rel err %.3e" % (E.item(), E_orc, eE, eg), flush=True)
    assert eE < 1e-13 and eg < 1e-10


def test_gradient_channel_matches_oracle(torch):
    print("\n[case] gradient, channel 16x12x20, 4 steps", flush=True)
    p = channel_params((16, 12, 20))
    u0 = list(cases.initial_state("channel", p, channel_pair(p)[0].grid, seed=6)[:3])
    gradient_vs_oracle(p, u0)


@pytest.mark.parametrize("kind", ["duct", "box"])
def test_gradient_walls_xyz_matches_oracle(torch, kind):
    print("\n[case] gradient, %s 16x12x20, 4 steps" % kind, flush=True)
    p = walled_params(kind, (16, 12, 20), is_forced=(1, 0, 0) if kind == "duct" else (0, 0, 0))
    rng = np.random.default_rng(7)
    u0 = [0.3 * rng.uniform(-1, 1, tuple(p["ng"])) for _ in range(3)]
    gradient_vs_oracle(p, u0)


@pytest.mark.timeout(600)
def test_channel_gradient_central_differences(torch):
    from cans_amd.adjoint import kinetic_energy_after_steps
    print("\n[case] channel 256x128x96 (tanh 1.8, Re_b 5600), 16 steps, checkpoints every 4", flush=True)
    p = dict(cases.params("channel"), ng=(256, 128, 96))
    sim, _ = channel_pair(p)
    u0 = list(cases.initial_state("channel", p, sim.grid, seed=3)[:3])
    u0[2][:, :, -1] = 0.
    dt, nsteps, every = 2e-4, 16, 4
    t0 = [torch.tensor(np.ascontiguousarray(a.transpose(2, 1, 0)), device="cuda", requires_grad=True) for a in u0]
    E = kinetic_energy_after_steps(sim, *t0, dt, nsteps, every)
    E.backward()
    grad = [t.grad.detach() for t in t0]
    base = [t.detach() for t in t0]
    gen = torch.Generator(device="cuda").manual_seed(2025)
    eps = 1e-6
    with torch.no_grad():
        for r in range(4):
            d = [torch.randn(x.shape, generator=gen, device="cuda", dtype=torch.float64) for x in base]
            d[2][-1] = 0.  # w's wall face is not a degree of freedom
            ep = kinetic_energy_after_steps(sim, *(x + eps * y for x, y in zip(base, d)), dt, nsteps, every).item()
            em = kinetic_energy_after_steps(sim, *(x - eps * y for x, y in zip(base, d)), dt, nsteps, every).item()
            fd = (ep - em) / (2 * eps)
            ad = sum(float((g * y).sum()) for g, y in zip(grad, d))
            rel = abs(fd - ad) / abs(ad)
            print("direction %d: central difference %.12e, adjoint %.12e, rel %.2e" % (r, fd, ad, rel), flush=True)
            assert rel < 1e-5, (r, fd, ad)


def imp_params(ng, forced=False):
    """Tri-periodic Taylor-Green box with 3-D implicit diffusion (a Helmholtz solve per component and
    substep), optionally with constant-flow-rate forcing of u."""
    p = cases.params("tgv", ng=ng)
    p.update(visci=20.0, is_impdiff=True)
    if forced:
        p.update(is_forced=(1, 0, 0), velf=(0.3, 0.0, 0.0))
    return p


def test_substep_adjoint_implicit_diffusion_matches_oracle(torch):
    print("\n[case] cans_rk3_substep_adjoint, implicit diffusion, tri-periodic 16x24x10", flush=True)
    substep_vs_oracle(imp_params((16, 24, 10), forced=True), 23)


def test_gradient_implicit_diffusion_matches_oracle(torch):
    print("\n[case] gradient, implicit diffusion, tri-periodic 16^3, 4 steps", flush=True)
    p = imp_params((16, 16, 16), forced=True)
    u0 = list(cases.initial_state("tgv", p, channel_pair(p)[0].grid, seed=8)[:3])
    gradient_vs_oracle(p, u0, dt=2e-2, nsteps=4)


@pytest.mark.parametrize("kind", ["tgv", "channel"])
def test_gradient_mixed_matches_fp64_oracle(torch, kind):
    """CANS_MIXED: FP32 pressure storage and transforms in the forward and adjoint projections; the
    gradient against the FP64 oracle's (measured 1.2e-10 tgv, 4.0e-10 channel; bound 1e-7, the pressure's
    FP32 rounding reaching the velocity through the correction's small dt)."""
    from cans_amd import _lib as LB
    from cans_amd.adjoint import kinetic_energy_after_steps
    from cans_amd.host import Simulation
    from oracle import adjoint as A
    print("\n[case] gradient, mixed precision, %s, 4 steps" % kind, flush=True)
    if kind == "tgv":
        p = cases.params("tgv", ng=(16, 16, 16))
        p["visci"] = 100.0
        dt = 5e-3
    else:
        p = channel_params((16, 12, 20))
        dt = 2e-3
    sim = Simulation(cases.host_case(p), precision=LB.CANS_MIXED)
    orc = cases.oracle_case(p)
    u0 = list(cases.initial_state(kind, p, sim.grid, seed=9)[:3])
    A._clear_wall_faces(orc, u0)
    E_orc, g_orc = A.grad_ke_after_steps(orc, *u0, dt, 4, every=2)
    t0 = [torch.tensor(np.ascontiguousarray(a.transpose(2, 1, 0)), device="cuda", requires_grad=True) for a in u0]
    E = kinetic_energy_after_steps(sim, *t0, dt, 4, every=2)
    E.backward()
    g_gpu = [t.grad.cpu().numpy().transpose(2, 1, 0) for t in t0]
    eE = abs(E.item() - E_orc) / E_orc
    eg = relerr(g_gpu, g_orc)
    print("E %.15e (FP64 oracle %.15e, rel %.2e); gradient rel err %.3e" % (E.item(), E_orc, eE, eg), flush=True)
    assert eE < 1e-9 and eg < 1e-7


def fd_check(sim, u0, dt, nsteps, every, seed, ndir=3, zero_faces=None):
    """Directional derivatives of the GPU gradient against central differences of the GPU's own steps."""
    import torch
    from cans_amd.adjoint import kinetic_energy_after_steps
    t0 = [torch.tensor(np.ascontiguousarray(a.transpose(2, 1, 0)), device="cuda", requires_grad=True) for a in u0]
    E = kinetic_energy_after_steps(sim, *t0, dt, nsteps, every)
    E.backward()
    grad = [t.grad.detach() for t in t0]
    base = [t.detach() for t in t0]
    gen = torch.Generator(device="cuda").manual_seed(seed)
    eps = 1e-6
    with torch.no_grad():
        for r in range(ndir):
            d = [torch.randn(x.shape, generator=gen, device="cuda", dtype=torch.float64) for x in base]
            if zero_faces:
                zero_faces(d)
            ep = kinetic_energy_after_steps(sim, *(x + eps * y for x, y in zip(base, d)), dt, nsteps, every).item()
            em = kinetic_energy_after_steps(sim, *(x - eps * y for x, y in zip(base, d)), dt, nsteps, every).item()
            fd = (ep - em) / (2 * eps)
            ad = sum(float((g * y).sum()) for g, y in zip(grad, d))
            rel = abs(fd - ad) / abs(ad)
            print("direction %d: central difference %.12e, adjoint %.12e, rel %.2e" % (r, fd, ad, rel), flush=True)
            assert rel < 1e-5, (r, fd, ad)


@pytest.mark.timeout(600)
def test_lid_driven_box_gradient_central_differences(torch):
    """Walls on all six faces (DCT pressure solves along x and y, stretched z), a moving lid: 96^3, 8 steps."""
    from cans_amd.adjoint import _clear_wall_faces
    print("\n[case] closed box 96^3 with a moving lid, 8 steps, checkpoints every 4", flush=True)
    p = walled_params("box", (96, 96, 96), visci=1000.0)
    sim, orc = channel_pair(p)
    rng = np.random.default_rng(10)
    u0 = [0.2 * rng.uniform(-1, 1, tuple(p["ng"])) for _ in range(3)]
    from oracle import adjoint as A
    A._clear_wall_faces(orc, u0)
    fd_check(sim, u0, 1e-3, 8, 4, 2026, zero_faces=lambda d: _clear_wall_faces(sim, d))


@pytest.mark.timeout(600)
def test_implicit_diffusion_gradient_central_differences(torch):
    """3-D implicit diffusion on a forced tri-periodic box: 128^3, 8 steps at a diffusive dt."""
    print("\n[case] implicit diffusion 128^3, 8 steps, checkpoints every 4", flush=True)
    p = imp_params((128, 128, 128), forced=True)
    p["visci"] = 50.0
    sim, _ = channel_pair(p)
    u0 = list(cases.initial_state("tgv", p, sim.grid, seed=11)[:3])
    fd_check(sim, u0, 1e-2, 8, 4, 2027)


def scalar_params(ng):
    """The heated channel (walls in z, stretched; the first scalar buoyant with gravity along x and z,
    Dirichlet walls and a source; the second passive with Neumann walls), both scalars bulk-forced."""
    p = cases.params("rb_channel", ng=ng)
    p["scalars"] = (dict(p["scalars"][0], is_forced=True, scalf=0.4), p["scalars"][1])
    return p


def test_substep_adjoint_scalars_matches_oracle(torch):
    from oracle import adjoint as A
    from oracle import oracle as O
    print("\n[case] cans_rk3_substep_adjoint_scal, heated channel 16x12x20, two scalars, buoyancy", flush=True)
    p = scalar_params((16, 12, 20))
    sim, orc = channel_pair(p)
    rng = np.random.default_rng(24)

    def dofs():
        a = [rng.uniform(-1, 1, tuple(orc.n)) for _ in range(3)]
        A._clear_wall_faces(orc, a)
        return a

    q, ub, fb = dofs(), dofs(), dofs()
    qh = [O.field(orc.n) for _ in range(3)]
    for x, y in zip(qh, q):
        x[1:-1, 1:-1, 1:-1] = y
    orc.bounduvw(*qh)
    ns = len(orc.scalars)
    s_in = [rng.uniform(-1, 1, tuple(orc.n)) for _ in range(ns)]
    sh = []
    for i, a in enumerate(s_in):
        f = O.field(orc.n)
        f[1:-1, 1:-1, 1:-1] = a
        orc.boundp(f, cbc=orc.scalars[i]["cbc"], bc=orc.scalars[i]["bc"])
        sh.append(f)
    sb = [rng.uniform(-1, 1, tuple(orc.n)) for _ in range(ns)]
    fsb = [rng.uniform(-1, 1, tuple(orc.n)) for _ in range(ns)]

    def put(arrays, halo):
        out = []
        for a in arrays:
            f = sim.ctx.field()
            f.set(a if halo else np.pad(a, 1))
            out.append(f)
        return out

    fq, fub, ffb = put(qh, True), put(ub, False), put(fb, False)
    fs, fsbf, ffsb = put(sh, True), put(sb, False), put(fsb, False)
    fubo, ffbo = [sim.ctx.field() for _ in range(3)], [sim.ctx.field() for _ in range(3)]
    fsbo, ffsbo = [sim.ctx.field() for _ in range(ns)], [sim.ctx.field() for _ in range(ns)]
    pp = sim.ctx.field()
    dt, irk = 2e-3, 2
    sim.rk3_substep_adjoint(irk, dt, [f.ptr for f in fq], fub, ffb, fubo, ffbo, pp,
                            ([f.ptr for f in fs], fsbf, ffsb, fsbo, ffsbo))
    sim.ctx.sync()
    r_ub, r_fb, r_sb, r_fsb = A.substep_adjoint(orc, (qh, sh), irk - 1, dt, ub, fb, sb, fsb)
    errs = [relerr([f.interior() for f in fubo], r_ub), relerr([f.interior() for f in ffbo], r_fb),
            relerr([f.interior() for f in fsbo], r_sb), relerr([f.interior() for f in ffsbo], r_fsb)]
    print("adjoint substep rel err: velocity %.3e, rhs %.3e, scalars %.3e, scalar rhs %.3e" % tuple(errs), flush=True)
    assert max(errs) < 1e-12


def test_gradient_scalars_matches_oracle(torch):
    from cans_amd.adjoint import kinetic_energy_after_steps
    from oracle import adjoint as A
    print("\n[case] gradient, heated channel 16x12x20 with two scalars and buoyancy, 4 steps", flush=True)
    p = scalar_params((16, 12, 20))
    sim, orc = channel_pair(p)
    u0 = list(cases.initial_state("rb_channel", p, sim.grid, seed=12)[:3])
    A._clear_wall_faces(orc, u0)
    s0 = cases.initial_scalars("rb_channel", p, sim.grid, seed=13)
    dt, nsteps = 2e-3, 4
    E_orc, g_orc = A.grad_ke_after_steps(orc, *u0, dt, nsteps, every=2, scal0=s0)
    t0 = [torch.tensor(np.ascontiguousarray(a.transpose(2, 1, 0)), device="cuda", requires_grad=True) for a in u0]
    ts = [torch.tensor(np.ascontiguousarray(a.transpose(2, 1, 0)), device="cuda") for a in s0]
    E = kinetic_energy_after_steps(sim, *t0, dt, nsteps, every=2, s0=ts)
    E.backward()
    g_gpu = [t.grad.cpu().numpy().transpose(2, 1, 0) for t in t0]
    eE = abs(E.item() - E_orc) / E_orc
    eg = relerr(g_gpu, g_orc)
    print("E %.15e (oracle %.15e, rel %.2e); gradient rel err %.3e" % (E.item(), E_orc, eE, eg), flush=True)
    assert eE < 1e-13 and eg < 1e-10


@pytest.mark.timeout(600)
def test_scalars_gradient_central_differences(torch):
    """The heated channel with two scalars and buoyancy at 128x64x96, 8 steps: central differences."""
    from cans_amd.adjoint import kinetic_energy_after_steps
    from oracle import adjoint as A
    print("\n[case] heated channel 128x64x96 with two scalars and buoyancy, 8 steps, checkpoints every 4", flush=True)
    p = scalar_params((128, 64, 96))
    sim, orc = channel_pair(p)
    u0 = list(cases.initial_state("rb_channel", p, sim.grid, seed=14)[:3])
    A._clear_wall_faces(orc, u0)
    s0 = [torch.tensor(np.ascontiguousarray(a.transpose(2, 1, 0)), device="cuda")
          for a in cases.initial_scalars("rb_channel", p, sim.grid, seed=15)]
    dt, nsteps, every = 1e-3, 8, 4
    t0 = [torch.tensor(np.ascontiguousarray(a.transpose(2, 1, 0)), device="cuda", requires_grad=True) for a in u0]
    E = kinetic_energy_after_steps(sim, *t0, dt, nsteps, every, s0=s0)
    E.backward()
    grad = [t.grad.detach() for t in t0]
    base = [t.detach() for t in t0]
    gen = torch.Generator(device="cuda").manual_seed(2028)
    eps = 1e-6
    with torch.no_grad():
        for r in range(3):
            d = [torch.randn(x.shape, generator=gen, device="cuda", dtype=torch.float64) for x in base]
            d[2][-1] = 0.
            ep = kinetic_energy_after_steps(sim, *(x + eps * y for x, y in zip(base, d)), dt, nsteps, every,
                                            s0=s0).item()
            em = kinetic_energy_after_steps(sim, *(x - eps * y for x, y in zip(base, d)), dt, nsteps, every,
                                            s0=s0).item()
            fd = (ep - em) / (2 * eps)
            ad = sum(float((g * y).sum()) for g, y in zip(grad, d))
            rel = abs(fd - ad) / abs(ad)
            print("direction %d: central difference %.12e, adjoint %.12e, rel %.2e" % (r, fd, ad, rel), flush=True)
            assert rel < 1e-5, (r, fd, ad)


def heated_cavity_params(ng):
    """Walls in x and z (y periodic), stretched z: the buoyant scalar Dirichlet on the x walls (hot, cold),
    Neumann in z; a passive one the other way round."""
    return dict(ng=ng, l=(1.0, 0.5, 1.0), gtype=1, gr=1.2, cbcvel="DDPPDD" * 3, cbcpre="NNPPNN", visci=60.0,
                scalars=(dict(alphai=80.0, cbc="DDPPNN", bc=(-0.5, 0.5, 0.0, 0.0, 0.1, 0.2)),
                         dict(alphai=120.0, cbc="NNPPDD", bc=(0.3, -0.2, 0.0, 0.0, 1.0, 0.0))),
                gacc=(0.0, 0.0, -1.0), beta=3.0, is_boussinesq_buoyancy=True)


def test_substep_adjoint_scalars_walls_matches_oracle(torch):
    from oracle import adjoint as A
    from oracle import oracle as O
    print("\n[case] cans_rk3_substep_adjoint_scal, heated cavity 16x8x20 (walls in x and z)", flush=True)
    p = heated_cavity_params((16, 8, 20))
    sim, orc = channel_pair(p)
    rng = np.random.default_rng(25)

    def dofs():
        a = [rng.uniform(-1, 1, tuple(orc.n)) for _ in range(3)]
        A._clear_wall_faces(orc, a)
        return a

    q, ub, fb = dofs(), dofs(), dofs()
    qh = [O.field(orc.n) for _ in range(3)]
    for x, y in zip(qh, q):
        x[1:-1, 1:-1, 1:-1] = y
    orc.bounduvw(*qh)
    ns = len(orc.scalars)
    sh = []
    for i in range(ns):
        f = O.field(orc.n)
        f[1:-1, 1:-1, 1:-1] = rng.uniform(-1, 1, tuple(orc.n))
        orc.boundp(f, cbc=orc.scalars[i]["cbc"], bc=orc.scalars[i]["bc"])
        sh.append(f)
    sb = [rng.uniform(-1, 1, tuple(orc.n)) for _ in range(ns)]
    fsb = [rng.uniform(-1, 1, tuple(orc.n)) for _ in range(ns)]

    def put(arrays, halo):
        out = []
        for a in arrays:
            f = sim.ctx.field()
            f.set(a if halo else np.pad(a, 1))
            out.append(f)
        return out

    fq, fub, ffb = put(qh, True), put(ub, False), put(fb, False)
    fs, fsbf, ffsb = put(sh, True), put(sb, False), put(fsb, False)
    fubo, ffbo = [sim.ctx.field() for _ in range(3)], [sim.ctx.field() for _ in range(3)]
    fsbo, ffsbo = [sim.ctx.field() for _ in range(ns)], [sim.ctx.field() for _ in range(ns)]
    pp = sim.ctx.field()
    dt, irk = 2e-3, 2
    sim.rk3_substep_adjoint(irk, dt, [f.ptr for f in fq], fub, ffb, fubo, ffbo, pp,
                            ([f.ptr for f in fs], fsbf, ffsb, fsbo, ffsbo))
    sim.ctx.sync()
    r_ub, r_fb, r_sb, r_fsb = A.substep_adjoint(orc, (qh, sh), irk - 1, dt, ub, fb, sb, fsb)
    errs = [relerr([f.interior() for f in fubo], r_ub), relerr([f.interior() for f in ffbo], r_fb),
            relerr([f.interior() for f in fsbo], r_sb), relerr([f.interior() for f in ffsbo], r_fsb)]
    print("adjoint substep rel err: velocity %.3e, rhs %.3e, scalars %.3e, scalar rhs %.3e" % tuple(errs), flush=True)
    assert max(errs) < 1e-12


def test_gradient_differentially_heated_cavity_matches_oracle(torch):
    """The reference's differentially heated cavity (tests/differentially_heated_cavity: walls in x and z, hot
    and cold x walls, buoyancy), at 64x2x64 from its linear temperature profile: the gradient after 4 steps."""
    from cans_amd.adjoint import kinetic_energy_after_steps
    from oracle import adjoint as A
    print("\n[case] gradient, differentially heated cavity 64x2x64, 4 steps", flush=True)
    p = cases.params("dhc", ng=(64, 2, 64))
    sim, orc = channel_pair(p)
    rng = np.random.default_rng(16)
    u0 = [0.05 * rng.uniform(-1, 1, tuple(p["ng"])) for _ in range(3)]
    A._clear_wall_faces(orc, u0)
    s0 = cases.initial_scalars("dhc", p, sim.grid)
    dt, nsteps = 2e-3, 4
    E_orc, g_orc = A.grad_ke_after_steps(orc, *u0, dt, nsteps, every=2, scal0=s0)
    t0 = [torch.tensor(np.ascontiguousarray(a.transpose(2, 1, 0)), device="cuda", requires_grad=True) for a in u0]
    ts = [torch.tensor(np.ascontiguousarray(a.transpose(2, 1, 0)), device="cuda") for a in s0]
    E = kinetic_energy_after_steps(sim, *t0, dt, nsteps, every=2, s0=ts)
    E.backward()
    g_gpu = [t.grad.cpu().numpy().transpose(2, 1, 0) for t in t0]
    eE = abs(E.item() - E_orc) / E_orc
    eg = relerr(g_gpu, g_orc)
    print("E %.15e (oracle %.15e, rel %.2e); gradient rel err %.3e" % (E.item(), E_orc, eE, eg), flush=True)
    assert eE < 1e-13 and eg < 1e-10


def test_adjoint_refusals(torch):
    """Out-of-scope substeps are refused with an error and leave the fields untouched: a step with scalars
    through the velocity-only entry point, implicit diffusion next to walls, z-only implicit diffusion."""
    from cans_amd import _lib as LB
    print("\n[case] adjoint refusals", flush=True)

    def attempt(p):
        sim, _ = channel_pair(p)
        f = [sim.ctx.field() for _ in range(12)]
        for x in f:
            x.buf.fill_(0.25)
        pp = sim.ctx.field(kind=LB.CANS_FIELD_PRESSURE)
        q = [sim.u.ptr, sim.v.ptr, sim.w.ptr]
        with pytest.raises(LB.CansError) as e:
            sim.rk3_substep_adjoint(2, 1e-3, q, f[0:3], f[3:6], f[6:9], f[9:12], pp)
        assert all(float(x.buf.min()) == 0.25 == float(x.buf.max()) for x in f[6:12])
        return str(e.value)

    msg = attempt(scalar_params((16, 12, 20)))
    assert "cans_rk3_substep_adjoint_scal" in msg
    p = channel_params((16, 12, 20))
    p["is_impdiff"] = True
    assert "tri-periodic" in attempt(p)
    p = channel_params((16, 12, 20))
    p.update(is_impdiff=True, is_impdiff_1d=True)
    assert "z-only" in attempt(p)
