"""Pins the oracle's scalar transport (scal.f90, rk_scal) and Boussinesq buoyancy.

Checks come from the mathematics, not from the routines' own formulas:
* the advective form is conservative and skew-symmetric, so for a discretely
  divergence-free velocity on a periodic box sum(dsdt) = 0 and sum(s*dsdt) = 0;
* a uniform velocity advects cos(kx) with the closed-form centred derivative;
* pure diffusion of cos(kx) has the eigenvalue -4 sin^2(k dx/2)/dx^2, and of zc^2
  on a stretched z grid the value alpha*(dzc(k) + dzc(k-1))*dzfi(k);
* rk_scal's update of a constant field, its source and its bulk forcing;
* the buoyancy term enters u, v, w at the faces between s(i) and s(i+1) etc.;
* the reference's own known-answer test tests/differentially_heated_cavity
  (Nusselt 8.8252 at step 10000, rtol/atol 1e-2, test.py:20).
"""
import numpy as np
import pytest

from oracle import oracle as O
from tests import cases


def periodic_case(ng=(16, 12, 10), l=(2.0, 1.5, 1.0), alphai=50.0):
    return O.OracleCase(ng=ng, l=l, gtype=1, gr=0.0, cbcvel="PP" * 9, cbcpre="PP" * 3, visci=100.0,
                        scalars=[dict(alphai=alphai, cbc="PP" * 3)])


def halo_periodic(a):
    """Interior (n1,n2,n3) array -> halo'd array with periodic ghost cells."""
    return np.asfortranarray(np.pad(a, 1, mode="wrap"))


def divfree_velocity(case, rng):
    """u = d(psi)/dy, v = -d(psi)/dx from a node-centred stream function: discretely divergence free."""
    n1, n2, n3 = case.n
    psi = rng.uniform(-1, 1, (n1, n2, n3))
    dxi, dyi = case.dli[0], case.dli[1]
    # psi at (i+1/2, j+1/2): u(i+1/2, j) = (psi(i,j) - psi(i,j-1))*dyi, v(i, j+1/2) = -(psi(i,j) - psi(i-1,j))*dxi
    u = (psi - np.roll(psi, 1, axis=1)) * dyi
    v = -(psi - np.roll(psi, 1, axis=0)) * dxi
    w = np.zeros_like(u)
    return halo_periodic(u), halo_periodic(v), halo_periodic(w)


def div_interior(case, u, v, w):
    dxi, dyi = case.dli[0], case.dli[1]
    dzfi = case.dzfi[1:-1]
    return ((u[1:-1, 1:-1, 1:-1] - u[:-2, 1:-1, 1:-1]) * dxi + (v[1:-1, 1:-1, 1:-1] - v[1:-1, :-2, 1:-1]) * dyi +
            (w[1:-1, 1:-1, 1:-1] - w[1:-1, 1:-1, :-2]) * dzfi[None, None, :])


def test_advection_conservative_and_skew():
    case = periodic_case()
    rng = np.random.default_rng(5)
    u, v, w = divfree_velocity(case, rng)
    assert np.abs(div_interior(case, u, v, w)).max() < 1e-12
    s = halo_periodic(rng.uniform(-1, 1, tuple(case.n)))
    dsdt = case.scal(u, v, w, s, 0.0)
    vol = np.prod(case.dl)
    assert abs(dsdt.sum() * vol) < 1e-12 * np.abs(dsdt).sum() * vol
    energy = (s[1:-1, 1:-1, 1:-1] * dsdt).sum()
    assert abs(energy) < 1e-12 * np.abs(s[1:-1, 1:-1, 1:-1] * dsdt).sum()
    # a constant scalar is not changed by a divergence-free velocity
    one = halo_periodic(np.ones(tuple(case.n)))
    assert np.abs(case.scal(u, v, w, one, 0.0)).max() < 1e-12


def test_uniform_advection_closed_form():
    case = periodic_case(ng=(32, 4, 4), l=(2 * np.pi, 1.0, 1.0))
    n1 = case.n[0]
    dx = case.dl[0]
    U = 0.7
    x = (np.arange(1, n1 + 1) - 0.5) * dx
    m = 3
    s = halo_periodic(np.broadcast_to(np.cos(m * x)[:, None, None], tuple(case.n)).copy())
    u = halo_periodic(np.full(tuple(case.n), U))
    zero = halo_periodic(np.zeros(tuple(case.n)))
    dsdt = case.scal(u, zero, zero, s, 0.0)
    # -U (s(i+1) - s(i-1)) / (2 dx) = U sin(m x) sin(m dx)/dx
    exact = U * np.sin(m * x) * np.sin(m * dx) / dx
    np.testing.assert_allclose(dsdt[:, 0, 0], exact, rtol=0, atol=1e-12)


def test_diffusion_closed_forms():
    alphai = 40.0
    case = periodic_case(ng=(24, 16, 8), alphai=alphai)
    alpha = 1.0 / alphai
    dx = case.dl[0]
    x = (np.arange(1, case.n[0] + 1) - 0.5) * dx
    m = 5
    kx = 2 * np.pi * m / case.l[0]
    s = halo_periodic(np.broadcast_to(np.cos(kx * x)[:, None, None], tuple(case.n)).copy())
    zero = halo_periodic(np.zeros(tuple(case.n)))
    dsdt = case.scal(zero, zero, zero, s, alpha)
    lam = -4.0 * np.sin(kx * dx / 2) ** 2 / dx ** 2
    np.testing.assert_allclose(dsdt[:, 3, 2], alpha * lam * np.cos(kx * x), rtol=0, atol=1e-10)
    # stretched z with walls: s = zc^2 gives alpha*(dzc(k) + dzc(k-1))*dzfi(k) in the interior rows
    sc = O.OracleCase(ng=(4, 4, 20), l=(1.0, 1.0, 1.0), gtype=1, gr=1.5, cbcvel="PPPPDD" * 3, cbcpre="PPPPNN",
                      visci=100.0, scalars=[dict(alphai=alphai, cbc="PPPPNN")])
    zc = sc.zc
    s2 = np.asfortranarray(np.broadcast_to((zc ** 2)[None, None, :], (6, 6, 22)).copy())
    z2 = np.zeros((6, 6, 22), order="F")
    d2 = sc.scal(z2, z2, z2, s2, alpha)
    k = np.arange(1, 21)
    exact = alpha * (sc.dzc[k] + sc.dzc[k - 1]) * sc.dzfi[k]
    np.testing.assert_allclose(d2[1, 2, :], exact, rtol=1e-11, atol=0)


def test_rk_scal_update_source_forcing():
    case = O.OracleCase(ng=(8, 6, 10), l=(1.0, 1.0, 1.0), gtype=1, gr=1.0, cbcvel="PPPPDD" * 3, cbcpre="PPPPNN",
                        visci=100.0, scalars=[dict(alphai=30.0, cbc="PPPPNN", source=0.25, is_forced=True,
                                                   scalf=2.0)])
    sc = case.scalars[0]
    c0 = 1.5
    s = np.full((10, 8, 12), c0, order="F")
    zero = np.zeros((10, 8, 12), order="F")
    rng = np.random.default_rng(3)
    ro = np.asfortranarray(rng.uniform(-1, 1, (8, 6, 10)))
    ro0 = ro.copy()
    rkpar = O.RKCOEFF[1]
    dt = 0.01
    f = case.rk_scal(rkpar, dt, sc, zero, zero, zero, s, ro)
    f1, f2 = rkpar[0] * dt, rkpar[1] * dt
    expect = c0 + f2 * ro0 + (f1 + f2) * 0.25  # a constant field has no transport
    np.testing.assert_allclose(s[1:-1, 1:-1, 1:-1], expect, rtol=1e-15, atol=1e-15)
    assert np.abs(ro).max() == 0.0  # dsdtrko <- dsdt = 0
    mean = (s[1:-1, 1:-1, 1:-1] * case.grid_vol_ratio_f[None, None, 1:-1]).sum()
    assert abs(f - (2.0 - mean)) < 1e-13
    case.bulk_forcing_s(1, f, s)
    mean2 = (s[1:-1, 1:-1, 1:-1] * case.grid_vol_ratio_f[None, None, 1:-1]).sum()
    assert abs(mean2 - 2.0) < 1e-13


def test_buoyancy_term():
    case = O.OracleCase(ng=(8, 6, 10), l=(1.0, 1.0, 1.0), gtype=1, gr=0.0, cbcvel="PP" * 9, cbcpre="PP" * 3,
                        visci=100.0, scalars=[dict(alphai=30.0, cbc="PP" * 3)], gacc=(0.3, -0.2, -1.0), beta=2.0,
                        is_boussinesq_buoyancy=True)
    st = case.new_state()
    rng = np.random.default_rng(9)
    s = halo_periodic(rng.uniform(-1, 1, (8, 6, 10)))
    rkpar = O.RKCOEFF[0]
    dt = 0.02
    case.rk(rkpar, dt, st["p"], st["dudtrko"], st["dvdtrko"], st["dwdtrko"], st["u"], st["v"], st["w"], s)
    f12 = (rkpar[0] + rkpar[1]) * dt
    si = s[1:-1, 1:-1, 1:-1]
    for q, d, g in ((st["u"], 0, 0.3), (st["v"], 1, -0.2), (st["w"], 2, -1.0)):
        nb = np.roll(s, -1, axis=d)[1:-1, 1:-1, 1:-1]  # s(i+1), s(j+1), s(k+1)
        np.testing.assert_allclose(q[1:-1, 1:-1, 1:-1], -f12 * g * 2.0 * 0.5 * (nb + si), rtol=1e-14, atol=1e-16)


@pytest.mark.timeout(600)
def test_dhc_nusselt_fixture():
    """tests/differentially_heated_cavity: 10000 adaptive-dt steps (cfl 0.95, dt every icheck = 10)."""
    p = cases.params("dhc")
    case = cases.oracle_case(p)
    st = case.new_state()
    st["s0"][1:-1, 1:-1, 1:-1] = cases.initial_scalars("dhc", p, case)[0]
    case.init_bounds(st)
    cfl, dtmax = 0.95, 1e9
    dt = min(cfl * case.chkdt(st["u"], st["v"], st["w"]), dtmax)
    for istep in range(1, 10001):
        case.rk3_step(st, dt)
        if istep % 10 == 0:
            dt = min(cfl * case.chkdt(st["u"], st["v"], st["w"]), dtmax)
    _, divmax = case.chkdiv(st["u"], st["v"], st["w"])
    assert divmax < 2.2e-9
    nu = cases.dhc_nusselt(st["s0"][1:-1, 1:-1, 1:-1], case.dl)
    np.testing.assert_allclose([nu], [8.8252], rtol=1e-2, atol=1e-2)


def test_gaussel_z_dense():
    """solver_gaussel_z (gaussel's eigenvalue-free branch) against dense solves of the same tridiagonal
    systems, for a c-grid system and an f-grid one with its last row dropped (q = 1)."""
    rng = np.random.default_rng(21)
    n = (3, 2, 12)
    for cf, bcz in (("c", "DD"), ("f", "DD"), ("c", "NN")):
        n3 = n[2]
        q = 1 if (cf == "f" and bcz[1] == "D") else 0
        nn = n3 - q
        a = rng.uniform(0.5, 1.0, n3)
        c = rng.uniform(0.5, 1.0, n3)
        b = -(a + c) - rng.uniform(1.0, 2.0, n3)  # diagonally dominant, like b + 1/alpha with alpha < 0
        norm = 0.37
        p = np.asfortranarray(rng.uniform(-1, 1, (n[0] + 2, n[1] + 2, n3 + 2)))
        ref = p.copy(order="F")
        A = np.diag(b[:nn]) + np.diag(a[1:nn], -1) + np.diag(c[:nn - 1], 1)
        for i in range(1, n[0] + 1):
            for j in range(1, n[1] + 1):
                ref[i, j, 1:nn + 1] = np.linalg.solve(A, p[i, j, 1:nn + 1] * norm)
        nn_, _k = O.IP(list(n))
        rc = O.lib().orc_solver_gaussel_z(nn_, O.P(a), O.P(b), O.P(c), bcz.encode(), cf.encode(), norm, O.P(p))
        assert rc == 0
        np.testing.assert_allclose(p, ref, rtol=1e-13, atol=1e-14)
