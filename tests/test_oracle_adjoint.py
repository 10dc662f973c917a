"""Pins the oracle's discrete adjoint of the RK3 step (oracle/adjoint.py).

* J^T of mom_xyz_ad by dot products: F is quadratic in the velocity, so
  (F(q + d) - F(q - d))/2 = J(q) d exactly and <g, J d> = <J^T g, d> to round-off,
  on a box with different spacings in x, y and z;
* the projection P of fillps/solver/correc: symmetric, idempotent, divergence-free;
* the gradient of the kinetic energy after a few RK3 steps (with checkpoints and
  recompute) against central differences of the oracle's own steps along random
  directions, and identical with and without checkpoints.
"""
import numpy as np
import pytest

from oracle import adjoint as A
from oracle import oracle as O


def tgv_case(ng, l=(2 * np.pi,) * 3, visci=1600.0):
    return O.OracleCase(ng=ng, l=l, gtype=1, gr=0.0, cbcvel="PP" * 9, cbcpre="PP" * 3, visci=visci)


def tgv_state(case, rng, noise=1e-2):
    """Taylor-Green velocity on the staggered grid (initflow.f90:102-117) plus uniform noise."""
    n1, n2, n3 = case.n
    dx, dy, dz = case.dl
    xc = (np.arange(1, n1 + 1) - 0.5) * dx
    yc = (np.arange(1, n2 + 1) - 0.5) * dy
    zc = (np.arange(1, n3 + 1) - 0.5) * dz
    xf, yf = xc + 0.5 * dx, yc + 0.5 * dy
    u = np.sin(xf)[:, None, None] * np.cos(yc)[None, :, None] * np.cos(zc)[None, None, :]
    v = -np.cos(xc)[:, None, None] * np.sin(yf)[None, :, None] * np.cos(zc)[None, None, :]
    w = np.zeros((n1, n2, n3))
    return [np.asfortranarray(q + noise * rng.uniform(-1, 1, q.shape)) for q in (u, v, w)]


def dot(a, b):
    return sum(float(np.sum(x * y)) for x, y in zip(a, b))


def fd_close(fd, ad, grad, d, rtol=1e-6):
    """|fd - ad| within rtol of max(|ad|, 1e-3 |grad| |d|): the oracle's bulk means are OpenMP reductions whose
    summation order varies between runs, so a direction nearly orthogonal to the gradient would otherwise
    test the difference quotient's rounding rather than the adjoint."""
    norm = np.sqrt(sum(float(np.sum(g * g)) for g in grad) * sum(float(np.sum(x * x)) for x in d))
    return abs(fd - ad) <= rtol * max(abs(ad), 1e-3 * norm)


def test_mom_jt_dot_product():
    case = tgv_case((8, 6, 5), l=(1.0, 1.5, 0.7), visci=37.0)
    rng = np.random.default_rng(1)
    q = [rng.uniform(-1, 1, tuple(case.n)) for _ in range(3)]
    d = [rng.uniform(-1, 1, tuple(case.n)) for _ in range(3)]
    g = [rng.uniform(-1, 1, tuple(case.n)) for _ in range(3)]
    fp = A.mom(case, *(A.halo(case, a + b) for a, b in zip(q, d)))
    fm = A.mom(case, *(A.halo(case, a - b) for a, b in zip(q, d)))
    jd = [(a - b) / 2 for a, b in zip(fp, fm)]
    lhs, rhs = dot(g, jd), dot(A.mom_jt(case, q, g), d)
    scale = sum(float(np.abs(x).sum()) for x in jd)
    assert abs(lhs - rhs) <= 1e-13 * scale, (lhs, rhs)
    # a plausible slip (dropping the coefficient term) is caught
    jt_bad = [case.visc * A._lap(case, x) for x in g]
    assert abs(dot(jt_bad, d) - lhs) > 1e-3 * abs(lhs)


def test_projection_symmetric_idempotent():
    case = tgv_case((8, 8, 6), l=(1.0, 2.0, 0.5))
    rng = np.random.default_rng(2)
    a = [rng.uniform(-1, 1, tuple(case.n)) for _ in range(3)]
    b = [rng.uniform(-1, 1, tuple(case.n)) for _ in range(3)]
    pa, pb = A.project(case, a), A.project(case, b)
    assert abs(dot(pa, b) - dot(a, pb)) <= 1e-13 * abs(dot(pa, b)) + 1e-14
    ppa = A.project(case, pa)
    assert max(float(np.abs(x - y).max()) for x, y in zip(ppa, pa)) < 1e-13
    _, divmax = case.chkdiv(*(A.halo(case, x) for x in pa))
    assert divmax < 1e-12


@pytest.mark.timeout(300)
def test_gradient_matches_central_differences():
    case = tgv_case((12, 12, 12), visci=100.0)
    rng = np.random.default_rng(3)
    u0 = tgv_state(case, rng)
    dt, nsteps = 5e-3, 5
    E, grad = A.grad_ke_after_steps(case, *u0, dt, nsteps, every=2)
    assert E == A.ke_after_steps(case, *u0, dt, nsteps)
    _, grad_all = A.grad_ke_after_steps(case, *u0, dt, nsteps, every=nsteps)
    for a, b in zip(grad, grad_all):  # recompute reproduces the forward bit for bit
        np.testing.assert_array_equal(a, b)
    eps = 1e-6
    for _ in range(3):
        d = [rng.standard_normal(tuple(case.n)) for _ in range(3)]
        ep = A.ke_after_steps(case, *(a + eps * b for a, b in zip(u0, d)), dt, nsteps)
        em = A.ke_after_steps(case, *(a - eps * b for a, b in zip(u0, d)), dt, nsteps)
        fd = (ep - em) / (2 * eps)
        ad = dot(grad, d)
        assert abs(fd - ad) <= 1e-6 * abs(ad), (fd, ad)


# ---- walls in z on a stretched grid (the channel), bulk forcing ----

def channel_case(ng=(8, 6, 10), gr=1.8, cbcvel="PPPPDD" * 3, bcvel=None, is_forced=(1, 0, 0), visci=90.0):
    return O.OracleCase(ng=ng, l=(2.0, 1.2, 1.0), gtype=1, gr=gr, cbcvel=cbcvel, cbcpre="PPPPNN",
                        bcvel=bcvel, visci=visci, is_forced=is_forced, velf=(1.0, 0.0, 0.0))


def dofs(case, rng):
    """Random values on the degrees of freedom (the normal velocities' wall faces zero)."""
    a = [rng.uniform(-1, 1, tuple(case.n)) for _ in range(3)]
    A._clear_wall_faces(case, a)
    return a


def channel_state(case, rng, noise=0.1):
    """Laminar profile u = 6 z (1 - z) on the cell centres plus uniform noise (initflow.f90:44-63)."""
    zc = np.asarray(case.zc[1:case.n[2] + 1])
    u = np.broadcast_to(6 * zc * (1 - zc), tuple(case.n))
    q = [u, np.zeros(tuple(case.n)), np.zeros(tuple(case.n))]
    out = [np.asfortranarray(x + noise * rng.uniform(-1, 1, x.shape)) for x in q]
    out[2][:, :, -1] = 0.
    return out


def ghosts(case, a):
    """Interior arrays -> halo'd fields with the case's own ghost values (bounduvw, non-zero wall velocity)."""
    f = [O.field(case.n) for _ in range(3)]
    for x, y in zip(f, a):
        x[1:-1, 1:-1, 1:-1] = y
    case.bounduvw(*f)
    return f


def test_mom_jt_walls_dot_product():
    # stretched z, a moving top wall for u (Dirichlet 1) and a Neumann v at the bottom: both folds
    bc = np.zeros(18)
    bc[5] = 1.0
    bc[10] = 0.3
    case = channel_case(cbcvel="PPPPDD" + "PPPPND" + "PPPPDD", bcvel=bc, visci=37.0)
    rng = np.random.default_rng(11)
    q, d, g = dofs(case, rng), dofs(case, rng), dofs(case, rng)
    fp = A.mom(case, *ghosts(case, [a + b for a, b in zip(q, d)]))
    fm = A.mom(case, *ghosts(case, [a - b for a, b in zip(q, d)]))
    jd = [(a - b) / 2 for a, b in zip(fp, fm)]
    A._clear_wall_faces(case, jd)  # F on a normal wall face is overwritten by bound
    lhs = dot(g, jd)
    jt = A.mom_jt_walls(case, ghosts(case, q), g)
    assert np.all(jt[2][:, :, -1] == 0.)
    rhs = dot(jt, d)
    scale = sum(float(np.abs(x).sum()) for x in jd)
    assert abs(lhs - rhs) <= 1e-13 * scale, (lhs, rhs)
    # planted mistakes: the ghost fold dropped, rdzm and rdzp swapped, the uniform-grid formula
    for bad in (A.mom_jt_walls(case, ghosts(case, q), g, fold=False),
                A.mom_jt_walls(case, ghosts(case, q), g, swap_rdz=True)):
        assert abs(dot(bad, d) - lhs) > 1e-6 * scale
    assert abs(dot(A.mom_jt(case, q, g), d) - lhs) > 1e-6 * scale


def test_projection_transpose_walls():
    case = channel_case(ng=(8, 8, 12))
    rng = np.random.default_rng(12)
    a, b = dofs(case, rng), dofs(case, rng)
    pa = A.project_dofs(case, a)
    assert np.all(pa[2][:, :, -1] == 0.)
    lhs, rhs = dot(A.project_t(case, a), b), dot(a, A.project_dofs(case, b))
    assert abs(lhs - rhs) <= 1e-12 * abs(lhs), (lhs, rhs)
    # the volume weights matter on the stretched grid: P itself is not symmetric
    assert abs(dot(pa, b) - rhs) > 1e-6 * abs(rhs)
    ppa = A.project_dofs(case, pa)
    assert max(float(np.abs(x - y).max()) for x, y in zip(ppa, pa)) < 1e-12
    _, divmax = case.chkdiv(*A.dof_fields(case, pa))
    assert divmax < 1e-11


def test_forcing_transpose():
    case = channel_case(is_forced=(1, 1, 0))
    rng = np.random.default_rng(13)
    a, b = dofs(case, rng), dofs(case, rng)
    gvr = np.asarray(case.grid_vol_ratio_f[1:case.n[2] + 1])[None, None, :]
    fa = [x - float(np.sum(gvr * x)) if c < 2 else x for c, x in enumerate(a)]  # Jacobian of u + velf - <u>
    assert abs(dot(A.forcing_t(case, b), a) - dot(b, fa)) <= 1e-13 * sum(float(np.abs(x).sum()) for x in b)
    u = np.full(tuple(case.n), 0.7)
    assert np.abs(A.forcing_t(case, [u, u, u])[0] - (0.7 - 0.7 * gvr * u.size)).max() < 1e-15 * u.size


@pytest.mark.timeout(300)
def test_gradient_channel_matches_central_differences():
    case = channel_case(ng=(8, 6, 10), visci=60.0)
    rng = np.random.default_rng(14)
    u0 = channel_state(case, rng)
    dt, nsteps = 2e-3, 4
    E, grad = A.grad_ke_after_steps(case, *u0, dt, nsteps, every=2)
    # the oracle's bulk mean is an OpenMP reduction, so repeated forward runs agree to round-off only
    assert E == pytest.approx(A.ke_after_steps(case, *u0, dt, nsteps), rel=1e-13)
    assert np.all(grad[2][:, :, -1] == 0.)
    eps = 1e-6
    for _ in range(3):
        d = dofs(case, rng)
        ep = A.ke_after_steps(case, *(a + eps * b for a, b in zip(u0, d)), dt, nsteps)
        em = A.ke_after_steps(case, *(a - eps * b for a, b in zip(u0, d)), dt, nsteps)
        fd = (ep - em) / (2 * eps)
        ad = dot(grad, d)
        assert fd_close(fd, ad, grad, d), (fd, ad)


# ---- walls in x and y too (duct, closed box): DCT-based pressure solves ----

def walled_case(kind, ng=(8, 6, 10), gr=1.5, visci=37.0, is_forced=(0, 0, 0)):
    """duct: periodic x, walls in y and z; box: walls on all six faces (a moving lid: u = 1 at z = lz,
    a Neumann u on the y = 0 wall); both on a tanh-stretched z."""
    bc = np.zeros(18)
    bc[5] = 1.0
    if kind == "duct":
        cbcvel = "PPNDDD" + "PPDDDD" + "PPDDDD"
        cbcpre = "PPNNNN"
    else:
        cbcvel = "DDNDDD" + "DDDDDD" + "DDDDDD"
        cbcpre = "NNNNNN"
    return O.OracleCase(ng=ng, l=(1.3, 1.0, 1.0), gtype=1, gr=gr, cbcvel=cbcvel, cbcpre=cbcpre, bcvel=bc,
                        visci=visci, is_forced=is_forced, velf=(1.0, 0.0, 0.0))


@pytest.mark.parametrize("kind", ["duct", "box"])
def test_mom_jt_walls_xyz_dot_product(kind):
    case = walled_case(kind)
    A.check_case(case)
    rng = np.random.default_rng(31)
    q, d, g = dofs(case, rng), dofs(case, rng), dofs(case, rng)
    fp = A.mom(case, *ghosts(case, [a + b for a, b in zip(q, d)]))
    fm = A.mom(case, *ghosts(case, [a - b for a, b in zip(q, d)]))
    jd = [(a - b) / 2 for a, b in zip(fp, fm)]
    A._clear_wall_faces(case, jd)
    lhs = dot(g, jd)
    jt = A.mom_jt_walls(case, ghosts(case, q), g)
    rhs = dot(jt, d)
    scale = sum(float(np.abs(x).sum()) for x in jd)
    assert abs(lhs - rhs) <= 1e-13 * scale, (lhs, rhs)
    assert abs(dot(A.mom_jt_walls(case, ghosts(case, q), g, fold=False), d) - lhs) > 1e-6 * scale


@pytest.mark.parametrize("kind", ["duct", "box"])
def test_projection_transpose_walls_xyz(kind):
    case = walled_case(kind, ng=(8, 8, 12))
    rng = np.random.default_rng(32)
    a, b = dofs(case, rng), dofs(case, rng)
    pa = A.project_dofs(case, a)
    lhs, rhs = dot(A.project_t(case, a), b), dot(a, A.project_dofs(case, b))
    assert abs(lhs - rhs) <= 1e-12 * abs(lhs), (lhs, rhs)
    assert abs(dot(pa, b) - rhs) > 1e-6 * abs(rhs)
    _, divmax = case.chkdiv(*A.dof_fields(case, pa))
    assert divmax < 1e-11


@pytest.mark.timeout(300)
@pytest.mark.parametrize("kind", ["duct", "box"])
def test_gradient_walls_xyz_matches_central_differences(kind):
    case = walled_case(kind, ng=(8, 6, 10), visci=60.0, is_forced=(1, 0, 0) if kind == "duct" else (0, 0, 0))
    rng = np.random.default_rng(33)
    u0 = dofs(case, rng)
    u0 = [0.3 * x for x in u0]
    dt, nsteps = 2e-3, 4
    E, grad = A.grad_ke_after_steps(case, *u0, dt, nsteps, every=2)
    for m in range(3):  # the normal wall faces carry no gradient
        idx = [slice(None)] * 3
        idx[m] = -1
        if case.cbcpre[2 * m] != "P":
            assert np.all(grad[m][tuple(idx)] == 0.)
    eps = 1e-6
    for _ in range(3):
        d = dofs(case, rng)
        ep = A.ke_after_steps(case, *(a + eps * b for a, b in zip(u0, d)), dt, nsteps)
        em = A.ke_after_steps(case, *(a - eps * b for a, b in zip(u0, d)), dt, nsteps)
        fd = (ep - em) / (2 * eps)
        ad = dot(grad, d)
        assert fd_close(fd, ad, grad, d), (fd, ad)


@pytest.mark.timeout(300)
def test_gradient_implicit_diffusion_matches_central_differences():
    """3-D implicit diffusion (Helmholtz solves in every substep) on a tri-periodic box."""
    case = O.OracleCase(ng=(12, 10, 8), l=(2 * np.pi,) * 3, gtype=1, gr=0.0, cbcvel="PP" * 9, cbcpre="PP" * 3,
                        visci=20.0, is_impdiff=True, is_forced=(1, 0, 0), velf=(0.3, 0.0, 0.0))
    rng = np.random.default_rng(41)
    u0 = tgv_state(case, rng, noise=0.1)
    dt, nsteps = 2e-2, 4
    E, grad = A.grad_ke_after_steps(case, *u0, dt, nsteps, every=2)
    eps = 1e-6
    for _ in range(3):
        d = [rng.standard_normal(tuple(case.n)) for _ in range(3)]
        ep = A.ke_after_steps(case, *(a + eps * b for a, b in zip(u0, d)), dt, nsteps)
        em = A.ke_after_steps(case, *(a - eps * b for a, b in zip(u0, d)), dt, nsteps)
        fd = (ep - em) / (2 * eps)
        ad = dot(grad, d)
        assert fd_close(fd, ad, grad, d), (fd, ad)
    # planted mistake: the adjoint without its Helmholtz solves misses the central difference
    helm = A.helmholtz_inv
    try:
        A.helmholtz_inv = lambda c, a, dtrk: a
        _, bad = A.grad_ke_after_steps(case, *u0, dt, nsteps, every=2)
    finally:
        A.helmholtz_inv = helm
    assert abs(dot(bad, d) - fd) > 1e-4 * abs(fd)


# ---- scalars: transport and Boussinesq buoyancy ----

def scalar_channel(ng=(8, 6, 10), gr=1.5, visci=60.0):
    """Channel (walls in z, stretched) with a buoyant scalar (Dirichlet walls, hot bottom, bulk-forced) and a
    passive one (Neumann walls with a flux); gravity along z."""
    scal = (dict(alphai=80.0, cbc="PPPPDD", bc=(0.0, 0.0, 0.0, 0.0, 1.0, -0.5), is_forced=True, scalf=0.2),
            dict(alphai=120.0, cbc="PPPPNN", bc=(0.0, 0.0, 0.0, 0.0, 0.2, 0.3), is_forced=False))
    return O.OracleCase(ng=ng, l=(2.0, 1.2, 1.0), gtype=1, gr=gr, cbcvel="PPPPDD" * 3, cbcpre="PPPPNN", visci=visci,
                        is_forced=(1, 0, 0), velf=(1.0, 0.0, 0.0), scalars=scal, gacc=(0.0, 0.0, -1.0), beta=2.0,
                        is_boussinesq_buoyancy=True)


def scalar_ghosts(case, a, i):
    f = O.field(case.n)
    f[1:-1, 1:-1, 1:-1] = a
    sc = case.scalars[i]
    case.boundp(f, cbc=sc["cbc"], bc=sc["bc"])
    return f


def test_scal_jt_dot_product():
    case = scalar_channel()
    A.check_case(case)
    rng = np.random.default_rng(51)
    for i, sc in enumerate(case.scalars):
        q, dq = dofs(case, rng), dofs(case, rng)
        s0, ds = rng.uniform(-1, 1, tuple(case.n)), rng.uniform(-1, 1, tuple(case.n))
        sig = rng.uniform(-1, 1, tuple(case.n))
        fp = case.scal(*ghosts(case, [a + b for a, b in zip(q, dq)]), scalar_ghosts(case, s0 + ds, i), sc["alpha"])
        fm = case.scal(*ghosts(case, [a - b for a, b in zip(q, dq)]), scalar_ghosts(case, s0 - ds, i), sc["alpha"])
        lhs = float(np.sum(sig * (fp - fm) / 2))
        js, jq = A.scal_jt(case, ghosts(case, q), scalar_ghosts(case, s0, i), sig, sc["alpha"], sc["cbc"])
        rhs = float(np.sum(js * ds)) + dot(jq, dq)
        scale = float(np.abs(fp - fm).sum())
        assert abs(lhs - rhs) <= 1e-13 * scale, (i, lhs, rhs)
        # planted mistake: the ghost fold of the walls dropped
        bad = A._fold_scalar
        try:
            A._fold_scalar = lambda c, out, cbc: np.asfortranarray(out[:, :, 2:c.n[2] + 2])
            js2, _ = A.scal_jt(case, ghosts(case, q), scalar_ghosts(case, s0, i), sig, sc["alpha"], sc["cbc"])
        finally:
            A._fold_scalar = bad
        assert abs(float(np.sum(js2 * ds)) + dot(jq, dq) - lhs) > 1e-6 * scale


def test_buoyancy_transpose():
    case = scalar_channel()
    rng = np.random.default_rng(52)
    wb = dofs(case, rng)
    ds = rng.uniform(-1, 1, tuple(case.n))
    coef = [0.3, -0.7, 1.1]
    sc = case.scalars[0]
    # the linear part of the term: ghosts with zero boundary values
    saved = sc["bc"]
    try:
        sc["bc"] = np.zeros(6)
        sg = scalar_ghosts(case, ds, 0)
    finally:
        sc["bc"] = saved
    fwd = [-coef[c] * (np.roll(sg, -1, axis=c) + sg)[1:-1, 1:-1, 1:-1] for c in range(3)]
    A._clear_wall_faces(case, fwd)
    lhs = dot(wb, fwd)
    rhs = float(np.sum(A.buoyancy_t(case, wb, sc["cbc"], coef) * ds))
    assert abs(lhs - rhs) <= 1e-13 * sum(float(np.abs(x).sum()) for x in fwd), (lhs, rhs)


@pytest.mark.timeout(300)
def test_gradient_scalars_matches_central_differences():
    case = scalar_channel()
    rng = np.random.default_rng(53)
    u0 = channel_state(case, rng)
    s0 = [rng.uniform(-0.5, 0.5, tuple(case.n)) for _ in case.scalars]
    dt, nsteps = 2e-3, 4
    E, grad = A.grad_ke_after_steps(case, *u0, dt, nsteps, every=2, scal0=s0)
    eps = 1e-6
    for _ in range(3):
        d = dofs(case, rng)
        ep = A.ke_after_steps(case, *(a + eps * b for a, b in zip(u0, d)), dt, nsteps, scal0=s0)
        em = A.ke_after_steps(case, *(a - eps * b for a, b in zip(u0, d)), dt, nsteps, scal0=s0)
        fd = (ep - em) / (2 * eps)
        ad = dot(grad, d)
        assert fd_close(fd, ad, grad, d), (fd, ad)
    # planted mistake: the scalars' adjoint left out (their effect on the velocity through buoyancy)
    sj = A.scal_jt
    try:
        A.scal_jt = lambda c, qh, sh, sig, al, cbc: (np.zeros(tuple(c.n)), [np.zeros(tuple(c.n))] * 3)
        _, bad = A.grad_ke_after_steps(case, *u0, dt, nsteps, every=2, scal0=s0)
    finally:
        A.scal_jt = sj
    assert abs(dot(bad, d) - fd) > 1e-6 * abs(fd)


def heated_cavity(ng=(8, 4, 10), visci=60.0):
    """A differentially heated cavity like the reference's (walls in x and z, y periodic): the buoyant scalar
    Dirichlet on the x walls (hot, cold) and Neumann in z; a passive one the other way round; stretched z."""
    scal = (dict(alphai=80.0, cbc="DDPPNN", bc=(-0.5, 0.5, 0.0, 0.0, 0.1, 0.2), is_forced=False),
            dict(alphai=120.0, cbc="NNPPDD", bc=(0.3, -0.2, 0.0, 0.0, 1.0, 0.0), is_forced=False))
    return O.OracleCase(ng=ng, l=(1.0, 0.5, 1.0), gtype=1, gr=1.2, cbcvel="DDPPDD" * 3, cbcpre="NNPPNN",
                        visci=visci, scalars=scal, gacc=(0.0, 0.0, -1.0), beta=3.0, is_boussinesq_buoyancy=True)


def test_scal_jt_walls_xz_dot_product():
    case = heated_cavity()
    A.check_case(case)
    rng = np.random.default_rng(61)
    for i, sc in enumerate(case.scalars):
        q, dq = dofs(case, rng), dofs(case, rng)
        s0, ds = rng.uniform(-1, 1, tuple(case.n)), rng.uniform(-1, 1, tuple(case.n))
        sig = rng.uniform(-1, 1, tuple(case.n))
        fp = case.scal(*ghosts(case, [a + b for a, b in zip(q, dq)]), scalar_ghosts(case, s0 + ds, i), sc["alpha"])
        fm = case.scal(*ghosts(case, [a - b for a, b in zip(q, dq)]), scalar_ghosts(case, s0 - ds, i), sc["alpha"])
        lhs = float(np.sum(sig * (fp - fm) / 2))
        js, jq = A.scal_jt(case, ghosts(case, q), scalar_ghosts(case, s0, i), sig, sc["alpha"], sc["cbc"])
        rhs = float(np.sum(js * ds)) + dot(jq, dq)
        scale = float(np.abs(fp - fm).sum())
        assert abs(lhs - rhs) <= 1e-13 * scale, (i, lhs, rhs)


@pytest.mark.timeout(300)
def test_gradient_heated_cavity_matches_central_differences():
    case = heated_cavity()
    rng = np.random.default_rng(62)
    u0 = [0.3 * x for x in dofs(case, rng)]
    s0 = [rng.uniform(-0.5, 0.5, tuple(case.n)) for _ in case.scalars]
    dt, nsteps = 2e-3, 4
    E, grad = A.grad_ke_after_steps(case, *u0, dt, nsteps, every=2, scal0=s0)
    eps = 1e-6
    for _ in range(3):
        d = dofs(case, rng)
        ep = A.ke_after_steps(case, *(a + eps * b for a, b in zip(u0, d)), dt, nsteps, scal0=s0)
        em = A.ke_after_steps(case, *(a - eps * b for a, b in zip(u0, d)), dt, nsteps, scal0=s0)
        fd = (ep - em) / (2 * eps)
        ad = dot(grad, d)
        assert fd_close(fd, ad, grad, d), (fd, ad)
