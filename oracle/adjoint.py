"""Discrete adjoint of the RK3 pressure-correction step (TEST INFRASTRUCTURE ONLY).

Like the rest of ``oracle/``, this module is the checker for the HIP path: only
``tests/`` and ``bench.py``'s cpu_baseline leg may use it.

Scope: tri-periodic boxes on uniform grids, explicit diffusion, no forcing,
scalars or buoyancy (a Taylor-Green flow).  There the velocity map of RK3
substep irk (src/main.f90:437-473 with rk.f90:24-333, fillps, solver, correc) is

    u' = P(u + f1 F(u) + f2 Fo),      Fo' = F(u),

with F the momentum right-hand side of mom_xyz_ad (mom.f90:771-1019), f1 =
rkcoeff(1,irk) dt, f2 = rkcoeff(2,irk) dt (rk.f90:78-80) and P = I - G L^-1 D the
projection that fillps/solver/correc apply.  The rk's -dtrk G p term drops out:
the solver's eigenvalues are those of the same 7-point Laplacian L = D G, so
P G p = 0 and the pressure only follows the velocity.  On a uniform periodic
grid G = -D^T, so P is symmetric and the adjoint of a substep, given the
adjoints ub of u' and fb of Fo', is

    wb = P ub,   ub <- wb + J(u)^T (f1 wb + fb),   fb <- f2 wb,

where J(u) is the Jacobian of F at the substep's starting velocity.

J^T: every advection term of mom_xyz_ad has the form K(x) q_c(x +- e_d) with
K^c_d(x) = (q_d(x) + q_d(x + e_c))/4 (c the component, d the direction; e.g.
uuip = 0.25 (u + u(i+1)) u(i+1), vujm = 0.25 (v(j-1) + v(i+1,j-1)) u(j-1)), so

    F_c(x) = -sum_d h_d [K^c_d(x) q_c(x+e_d) - K^c_d(x-e_d) q_c(x-e_d)] + nu L q_c,

and with M^c_d(x) = g_c(x) q_c(x+e_d) - g_c(x+e_d) q_c(x),

    (J^T g)_c(y) = -sum_d h_d [K^c_d(y-e_d) g_c(y-e_d) - K^c_d(y) g_c(y+e_d)]
                   - h_c/4 sum_c' [M^c'_c(y) + M^c'_c(y-e_c')] + nu L g_c(y).

The tests pin mom_jt by dot products against central differences of the
oracle's own F (exact: F is quadratic), P by symmetry and idempotence, and the
whole gradient by central differences of the oracle's RK3 steps.
"""
import numpy as np

from oracle import oracle as O


def interior(a):
    return np.asfortranarray(a[1:-1, 1:-1, 1:-1])


def halo(case, a):
    """Interior (n1,n2,n3) -> halo'd array with periodic ghost cells."""
    return np.asfortranarray(np.pad(a, 1, mode="wrap"))


def check_case(case):
    if case.cbcpre != "PP" * 3 or case.cbcvel != "PP" * 9:
        raise ValueError("the adjoint covers tri-periodic boxes")
    if np.ptp(case.dzci) > 1e-12 * np.abs(case.dzci).max():
        raise ValueError("the adjoint covers uniform grids")
    if any(case.is_forced) or case.scalars or case.is_impdiff:
        raise ValueError("the adjoint covers unforced flows with explicit diffusion and no scalars")


def mom(case, u, v, w):
    """F = mom_xyz_ad(u, v, w) (explicit diffusion) as three interior arrays; u, v, w halo'd."""
    n, _k = case._n()
    out = [np.zeros(tuple(case.n), order="F") for _ in range(3)]
    O.lib().orc_mom_xyz_ad(n, case.dli[0], case.dli[1], O.P(case.dzci), O.P(case.dzfi), case.visc,
                           O.P(u), O.P(v), O.P(w), O.P(out[0]), O.P(out[1]), O.P(out[2]), None, None, None, 0, 0)
    return out


def _sh(a, d, s):
    """a(x + s e_d) on a periodic array."""
    return np.roll(a, -s, axis=d)


def _lap(case, a):
    h2 = [case.dli[0] ** 2, case.dli[1] ** 2, case.dzci[1] ** 2]
    return sum(h2[d] * (_sh(a, d, 1) - 2 * a + _sh(a, d, -1)) for d in range(3))


def mom_jt(case, q, g):
    """(dF/dq)^T g at velocity q; q and g are lists of three interior arrays."""
    h = [case.dli[0], case.dli[1], case.dzci[1]]
    out = [case.visc * _lap(case, gc) for gc in g]
    for c in range(3):
        for d in range(3):
            K = 0.25 * (q[d] + _sh(q[d], c, 1))
            M = g[c] * _sh(q[c], d, 1) - _sh(g[c], d, 1) * q[c]
            out[c] -= h[d] * (_sh(K * g[c], d, -1) - K * _sh(g[c], d, 1))
            out[d] -= 0.25 * h[d] * (M + _sh(M, c, -1))
    return out


def project(case, a):
    """P a through the oracle's fillps (dti = 1), solver and correc (dt = 1); a: three interior arrays."""
    u, v, w = (halo(case, x) for x in a)
    pp = O.field(case.n)
    case.fillps(u, v, w, pp, 1.0)
    case.solver(pp)
    case.boundp(pp)
    case.correc(pp, u, v, w, 1.0)
    return [interior(x) for x in (u, v, w)]


def kinetic_energy(u, v, w):
    """Volume-averaged kinetic energy of interior arrays on a uniform grid."""
    return 0.5 * float(np.mean(u * u) + np.mean(v * v) + np.mean(w * w))


def _state(case, u0, v0, w0):
    st = case.new_state()
    for k, q in zip("uvw", (u0, v0, w0)):
        st[k][...] = halo(case, q)
    case.init_bounds(st)
    return st


def _step_recording(case, st, dt):
    """One RK3 step (the oracle's rk3_step sequence) returning the interior velocity at the start of each
    substep."""
    u, v, w, p, pp = st["u"], st["v"], st["w"], st["p"], st["pp"]
    starts = []
    for irk in range(3):
        starts.append([interior(x) for x in (u, v, w)])
        rkpar = O.RKCOEFF[irk]
        dtrk = (rkpar[0] + rkpar[1]) * dt
        case.rk(rkpar, dt, p, st["dudtrko"], st["dvdtrko"], st["dwdtrko"], u, v, w)
        case.bounduvw(u, v, w, 0, -1)
        case.fillps(u, v, w, pp, dtrk ** (-1))
        case.updt_rhs_b(pp)
        case.solver(pp)
        case.boundp(pp)
        case.correc(pp, u, v, w, dtrk)
        case.bounduvw(u, v, w, 1, 1, 0)
        case.updatep(pp, p, 0.0)
        case.boundp(p, 1)
    return starts


def step_adjoint(case, starts, ub, dt):
    """Adjoint of one step: ub (adjoint of the step's final velocity) -> adjoint of its initial velocity."""
    fb = [np.zeros_like(x) for x in ub]
    for irk in (2, 1, 0):
        f1, f2 = O.RKCOEFF[irk][0] * dt, O.RKCOEFF[irk][1] * dt
        wb = project(case, ub)
        g = [f1 * a + b for a, b in zip(wb, fb)]
        fb = [f2 * a for a in wb]
        jt = mom_jt(case, starts[irk], g)
        ub = [a + b for a, b in zip(wb, jt)]
    return ub


def ke_after_steps(case, u0, v0, w0, dt, nsteps):
    """Kinetic energy after nsteps RK3 steps from the interior velocity (u0, v0, w0), zero pressure."""
    st = _state(case, u0, v0, w0)
    for _ in range(nsteps):
        case.rk3_step(st, dt)
    return kinetic_energy(*(interior(st[k]) for k in "uvw"))


def grad_ke_after_steps(case, u0, v0, w0, dt, nsteps, every=4):
    """(E, dE/du0): E the kinetic energy after nsteps steps; checkpoints of the velocity every `every`
    steps, the steps in between recomputed during the backward sweep."""
    check_case(case)
    st = _state(case, u0, v0, w0)
    ckpt = {}
    for s in range(nsteps):
        if s % every == 0:
            ckpt[s] = {k: st[k].copy(order="F") for k in st}
        case.rk3_step(st, dt)
    fin = [interior(st[k]) for k in "uvw"]
    E = kinetic_energy(*fin)
    ncell = float(np.prod(case.n))
    ub = [x / ncell for x in fin]
    for s0 in sorted(ckpt, reverse=True):
        seg = {k: v.copy(order="F") for k, v in ckpt[s0].items()}
        rec = [_step_recording(case, seg, dt) for _ in range(s0, min(s0 + every, nsteps))]
        for starts in reversed(rec):
            ub = step_adjoint(case, starts, ub, dt)
    return E, ub
