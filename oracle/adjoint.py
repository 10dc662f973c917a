"""Discrete adjoint of the RK3 pressure-correction step (TEST INFRASTRUCTURE ONLY).

Like the rest of ``oracle/``, this module is the checker for the HIP path: only
``tests/`` and ``bench.py``'s cpu_baseline leg may use it.

Scope: each direction either periodic or bounded by walls (Neumann pressure,
Dirichlet normal velocity, Dirichlet or Neumann tangential velocity); a periodic
z on a uniform grid, walls in z on any (stretched) grid; explicit diffusion, or
3-D implicit diffusion on a tri-periodic uniform grid (see substep_adjoint); bulk
forcing of u and v along periodic x and y; no scalars.  The velocity map of RK3
substep irk (src/main.f90:437-473 with rk.f90:24-333, fillps, solver, correc) on
the degrees of freedom (every interior value except a normal velocity's wall face,
u(n1), v(n2) or w(n3), which bound.f90 fixes) is

    u' = P Fr(u + f1 F(B u) + f2 Fo),      Fo' = F(B u),

with F the momentum right-hand side of mom_xyz_ad (mom.f90:771-1019) at the
velocity whose ghost cells B sets (bound.f90:17-122: affine in the interior),
f1 = rkcoeff(1,irk) dt, f2 = rkcoeff(2,irk) dt (rk.f90:78-80), Fr the bulk forcing
u <- u + (velf - sum_k gvr_f(k) u) (rk.f90:493-518, mom.f90:725-769) and
P = I - G L^-1 D the projection that fillps/solver/correc apply.  The rk's
-dtrk G p term drops out: the solver's eigenvalues and tridiagonal coefficients
are those of the same 7-point L = D G (homogeneous Neumann rows at the walls),
so P G p = 0 and the pressure only follows the velocity.

P is self-adjoint in the inner product weighted by the cell volumes M (dzf(k)
for u and v, dzc(k) for w; D and G are negative adjoints there:
sum_k dzf(k) p(k) (w(k) - w(k-1))/dzf(k) = -sum_k dzc(k) w(k) (p(k+1) - p(k))/dzc(k)
with w = 0 on the walls, and likewise along x and y with uniform weights), so its
transpose is P^T = M P M^-1; with a uniform z, M is a multiple of I and P^T = P.
This holds whatever transforms the solver uses along x and y: with walls there the
DCT pair REDFT10/REDFT01 (each the other's transpose up to scale) diagonalises the
same symmetric L, so the forward solve is its own weighted transpose.  The forcing's Jacobian
I - 1 gvr_f^T has the transpose I - gvr_f 1^T.  Given the adjoints ub of u' and
fb of Fo', the adjoint of a substep is

    wb = Fr'^T P^T ub,   ub <- wb + J(u)^T (f1 wb + fb),   fb <- f2 wb,

with J(u) the Jacobian of F(B u) at the substep's starting velocity.

J^T: every advection term of mom_xyz_ad has the form K(x) q_c(x +- e_d), with
K^c_d(x) = alpha^c_d(k) q_d(x) + beta^c_d(k) q_d(x + e_c): 1/4 and 1/4, except
for w advected along x and y on a stretched grid, where the reference weighs u
and v by rdzm(k)/4 = dzci(k)/dzfi(k)/4 and rdzp(k)/4 = dzci(k)/dzfi(k+1)/4
(e.g. uuip = 0.25 (u + u(i+1)) u(i+1), uwip = 0.25 (rdzm u + rdzp u(k+1)) w(i+1)),
so that, with H^c_d = dxi, dyi along x, y and dzfi(k) (u, v) or dzci(k) (w) along z,

    F_c(x) = -sum_d H^c_d(k) [K^c_d(x) q_c(x+e_d) - K^c_d(x-e_d) q_c(x-e_d)] + visc L_c q_c.

With G^c_d(x) = H^c_d(k_x) g_c(x) and M^c_d(x) = G^c_d(x) q_c(x+e_d) - G^c_d(x+e_d) q_c(x),

    (J^T g)_c(y) = -sum_d [K^c_d(y-e_d) G^c_d(y-e_d) - K^c_d(y) G^c_d(y+e_d)]
                   -sum_c' [alpha^c'_c(k_y) M^c'_c(y) + beta^c'_c(k_y - [c' = z]) M^c'_c(y-e_c')]
                   + visc L_c^T g_c(y),

evaluated at every cell and ghost cell y, g being zero wherever F is not a degree
of freedom; B^T then folds the ghost values into the interior (bound.f90:182-463:
u(0) = 2 ub - u(1) gives -1, a Neumann ghost +1; a normal velocity's wall faces
are constants).  bound.f90 sets the ghosts direction by direction, x, then y,
then z, each over the others' ghost layers too, so B^T folds z first, then y,
then x.

The tests pin mom_jt by dot products against central differences of the
oracle's own F (exact: F is quadratic) through the oracle's own ghost map, P^T by
<P^T a, b> = <a, P b>, and the whole gradient by central differences of the
oracle's RK3 steps.
"""
import numpy as np

from oracle import oracle as O


def interior(a):
    return np.asfortranarray(a[1:-1, 1:-1, 1:-1])


def halo(case, a):
    """Interior (n1,n2,n3) -> halo'd array with periodic ghost cells."""
    return np.asfortranarray(np.pad(a, 1, mode="wrap"))


def wall_dirs(case):
    """The directions bounded by walls (the pressure is Neumann there)."""
    return [d for d in range(3) if case.cbcpre[2 * d:2 * d + 2] != "PP"]


def walls(case):
    """True when some direction is bounded by walls."""
    return bool(wall_dirs(case))


def check_case(case):
    for d in range(3):
        cp = case.cbcpre[2 * d:2 * d + 2]
        cv = [case.cbcvel[6 * c + 2 * d:6 * c + 2 * d + 2] for c in range(3)]
        if cp == "PP":
            if any(x != "PP" for x in cv):
                raise ValueError("velocity and pressure disagree on a periodic direction")
        elif cp != "NN" or cv[d] != "DD" or any(ch not in "DN" for c in range(3) if c != d for ch in cv[c]):
            raise ValueError("walls: Neumann pressure, Dirichlet normal and Dirichlet or Neumann tangential velocity")
        elif case.n[d] < 2:
            raise ValueError("walls need at least two cells across")
    if 2 not in wall_dirs(case) and np.ptp(case.dzci) > 1e-12 * np.abs(case.dzci).max():
        raise ValueError("a periodic z needs a uniform grid")
    if case.is_forced[2] or any(case.is_forced[c] and c in wall_dirs(case) for c in (0, 1)):
        raise ValueError("forcing covers u and v along periodic directions")
    if case.scalars:
        if case.is_impdiff:
            raise ValueError("scalars: explicit diffusion")
        for sc in case.scalars:
            for d in range(3):
                cb = sc["cbc"][2 * d:2 * d + 2]
                if (d in wall_dirs(case)) != (cb != "PP") or any(ch not in "DNP" for ch in cb):
                    raise ValueError("scalars: periodic where the box is, Dirichlet or Neumann at its walls")
    if case.is_impdiff and (walls(case) or case.is_impdiff_1d):
        raise ValueError("implicit diffusion: the adjoint covers the 3-D form on tri-periodic uniform grids")


def _clear_wall_faces(case, a):
    """Zero each normal velocity's wall face (u(n1), v(n2), w(n3) on an interior array)."""
    for d in wall_dirs(case):
        idx = [slice(None)] * 3
        idx[d] = -1
        a[d][tuple(idx)] = 0.


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


def mom_jt(case, q, g, visc=None):
    """(dF/dq)^T g at velocity q; q and g are lists of three interior arrays (visc: the diffusion
    coefficient, case.visc unless given; 0 for the advection alone)."""
    h = [case.dli[0], case.dli[1], case.dzci[1]]
    visc = case.visc if visc is None else visc
    out = [visc * _lap(case, gc) for gc in g]
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


class _Homogeneous:
    """The case with zero boundary values: the linear part of bound.f90's ghost map."""

    def __init__(self, case):
        self.case = case

    def __enter__(self):
        self.saved = (self.case.bcvel, self.case.bcpre)
        self.case.bcvel = np.zeros_like(self.case.bcvel)
        self.case.bcpre = np.zeros_like(self.case.bcpre)
        return self.case

    def __exit__(self, *exc):
        self.case.bcvel, self.case.bcpre = self.saved


def dof_fields(case, a):
    """Three interior arrays -> halo'd fields with the homogeneous ghost values of bounduvw."""
    f = [O.field(case.n) for _ in range(3)]
    for x, y in zip(f, a):
        x[1:-1, 1:-1, 1:-1] = y
    for d in wall_dirs(case):  # the normal velocity's wall face
        idx = [slice(None)] * 3
        idx[d] = -2
        f[d][tuple(idx)] = 0.
    with _Homogeneous(case) as hc:
        hc.bounduvw(*f)
    return f


def project_dofs(case, a):
    """P a on the degrees of freedom (walls: homogeneous ghosts, the normal wall faces zero)."""
    u, v, w = dof_fields(case, a)
    pp = O.field(case.n)
    with _Homogeneous(case) as hc:
        hc.fillps(u, v, w, pp, 1.0)
        hc.solver(pp)
        hc.boundp(pp)
        hc.correc(pp, u, v, w, 1.0)
    return [interior(x) for x in (u, v, w)]


def volume_weights(case):
    """Volume weights of u, v and w as interior arrays: dzf, dzf, dzc along z; 0 on the normal wall faces."""
    n3 = case.n[2]
    wf = np.asarray(case.dzf[1:n3 + 1], dtype=np.float64)[None, None, :]
    wc = np.asarray(case.dzc[1:n3 + 1], dtype=np.float64)[None, None, :]
    m = [np.broadcast_to(x, tuple(case.n)).copy() for x in (wf, wf, wc)]
    _clear_wall_faces(case, m)
    return m


def project_t(case, a):
    """P^T a = M P M^-1 a (P^T = P with a uniform z, up to the wall faces, which it zeroes)."""
    if not walls(case):
        return project(case, a)
    m = volume_weights(case)
    safe = [np.where(x > 0, x, 1.) for x in m]
    pa = project_dofs(case, [y / s for y, s in zip(a, safe)])
    return [x * y for x, y in zip(pa, m)]


def forcing_t(case, b):
    """Transpose of the bulk forcing's Jacobian, I - gvr_f 1^T, on each forced component."""
    n3 = case.n[2]
    gvr = np.asarray(case.grid_vol_ratio_f[1:n3 + 1], dtype=np.float64)[None, None, :]
    out = [x.copy() for x in b]
    for c in (0, 1):
        if case.is_forced[c]:
            out[c] = out[c] - gvr * float(np.sum(b[c]))
    return out


def mom_jt_walls(case, qh, g, fold=True, swap_rdz=False):
    """(d F(B q)/dq)^T g on a box with walls: qh three halo'd fields (ghosts as the forward set them), g
    three interior arrays (ignored on the normal wall faces).  Returns three interior arrays, the normal
    wall faces zero.  fold=False and swap_rdz=True plant mistakes for the tests' negative controls."""
    n = list(case.n)
    wd = wall_dirs(case)
    # along a wall direction the arrays hold the ghost layers plus one zero layer on each side (index
    # i + 1 for i = -1 .. n + 2); along a periodic one the interior only, and np.roll wraps
    qsl = tuple(slice(None) if d in wd else slice(1, -1) for d in range(3))
    qpad = [(1, 1) if d in wd else (0, 0) for d in range(3)]
    gpad = [(2, 2) if d in wd else (0, 0) for d in range(3)]
    Q = [np.pad(x[qsl], qpad) for x in qh]
    G = [np.pad(x, gpad) for x in g]
    for d in wd:  # the normal velocity's wall face i = n_d is not a degree of freedom
        idx = [slice(None)] * 3
        idx[d] = n[d] + 1
        G[d][tuple(idx)] = 0.
    n3 = n[2]
    if 2 in wd:
        zc, zf = np.zeros(n3 + 4), np.zeros(n3 + 4)
        zc[1:n3 + 3], zf[1:n3 + 3] = case.dzci, case.dzfi
        rdzm, rdzp = np.zeros(n3 + 4), np.zeros(n3 + 4)
        rdzm[2:n3 + 2] = zc[2:n3 + 2] / zf[2:n3 + 2]
        rdzp[2:n3 + 2] = zc[2:n3 + 2] / zf[3:n3 + 3]
    else:
        zc, zf = np.asarray(case.dzci[1:n3 + 1], float), np.asarray(case.dzfi[1:n3 + 1], float)
        rdzm, rdzp = zc / zf, zc / np.roll(zf, -1)
    if swap_rdz:
        rdzm, rdzp = rdzp, rdzm
    zc_, zf_, rdzm_, rdzp_ = (x[None, None, :] for x in (zc, zf, rdzm, rdzp))
    h = case.dli[:2]

    def H(c, d):
        return h[d] if d < 2 else (zc_ if c == 2 else zf_)

    def al(c, d):
        return 0.25 * rdzm_ if (c == 2 and d < 2) else 0.25

    def be(c, d):
        return 0.25 * rdzp_ if (c == 2 and d < 2) else 0.25

    out = [np.zeros_like(x) for x in G]
    for m in range(3):  # visc L_m^T g_m
        gm = G[m]
        xy = sum(h[d] ** 2 * (_sh(gm, d, 1) - 2 * gm + _sh(gm, d, -1)) for d in range(2))
        if m < 2:
            gh = zf_ * gm
            z = zc_ * (_sh(gh, 2, 1) - gh) - _sh(zc_, 2, -1) * (gh - _sh(gh, 2, -1))
        else:
            gh = zc_ * gm
            z = _sh(zf_, 2, 1) * (_sh(gh, 2, 1) - gh) - zf_ * (gh - _sh(gh, 2, -1))
        out[m] += case.visc * (xy + z)
    for c in range(3):
        for d in range(3):
            K = al(c, d) * Q[d] + be(c, d) * _sh(Q[d], c, 1)
            Gd = H(c, d) * G[c]
            out[c] -= _sh(K * Gd, d, -1) - K * _sh(Gd, d, 1)
            M = Gd * _sh(Q[c], d, 1) - _sh(Gd, d, 1) * Q[c]
            out[d] -= al(c, d) * M + _sh(be(c, d) * M, c, -1)
    # B^T, the reverse of bound's x, y, z order: z first
    for d in sorted(wd, reverse=True) if fold else ():
        lo, hi, f_lo, f_hi, face = ([slice(None)] * 3 for _ in range(5))
        lo[d], hi[d], f_lo[d], f_hi[d], face[d] = 1, n[d] + 2, 2, n[d] + 1, n[d] + 1
        for m in range(3):
            if m == d:  # constants: the wall faces and whatever lies beyond them
                for sl in (lo, hi, face):
                    out[m][tuple(sl)] = 0.
                continue
            s0 = -1. if case.cbcvel[6 * m + 2 * d] == "D" else 1.
            s1 = -1. if case.cbcvel[6 * m + 2 * d + 1] == "D" else 1.
            out[m][tuple(f_lo)] += s0 * out[m][tuple(lo)]
            out[m][tuple(f_hi)] += s1 * out[m][tuple(hi)]
            out[m][tuple(lo)] = 0.
            out[m][tuple(hi)] = 0.
    isl = tuple(slice(2, n[d] + 2) if d in wd else slice(None) for d in range(3))
    res = [np.asfortranarray(x[isl]) for x in out]
    _clear_wall_faces(case, res)
    return res


def _ext(case, a, halo):
    """Along a wall direction: the ghost layers plus one zero layer each side (index i + 1 for
    i = -1 .. n + 2); along a periodic one the interior only (np.roll wraps).  halo: a is halo'd (its ghosts
    kept) or interior (its ghost layers zero)."""
    wd = wall_dirs(case)
    if halo:
        sl = tuple(slice(None) if d in wd else slice(1, -1) for d in range(3))
        return np.pad(a[sl], [(1, 1) if d in wd else (0, 0) for d in range(3)])
    return np.pad(a, [(2, 2) if d in wd else (0, 0) for d in range(3)])


def _z_tables(case):
    n3 = case.n[2]
    if 2 in wall_dirs(case):
        zc, zf = np.zeros(n3 + 4), np.zeros(n3 + 4)
        zc[1:n3 + 3], zf[1:n3 + 3] = case.dzci, case.dzfi
    else:
        zc, zf = np.asarray(case.dzci[1:n3 + 1], float), np.asarray(case.dzfi[1:n3 + 1], float)
    return zc[None, None, :], zf[None, None, :]


def _fold_scalar(case, out, cbc):
    """B^T of a cell-centred scalar's ghosts at the walls (s(0) = 2 sb - s(1): -1 Dirichlet; s(0) = s(1) -
    dz sb: +1 Neumann), z first, then y, then x (boundp sets them x, y, z); returns the interior."""
    out = out.copy()
    n = list(case.n)
    wd = wall_dirs(case)
    for d in sorted(wd, reverse=True):
        lo, hi, f_lo, f_hi = ([slice(None)] * 3 for _ in range(4))
        lo[d], hi[d], f_lo[d], f_hi[d] = 1, n[d] + 2, 2, n[d] + 1
        out[tuple(f_lo)] += (-1. if cbc[2 * d] == "D" else 1.) * out[tuple(lo)]
        out[tuple(f_hi)] += (-1. if cbc[2 * d + 1] == "D" else 1.) * out[tuple(hi)]
        out[tuple(lo)] = 0.
        out[tuple(hi)] = 0.
    isl = tuple(slice(2, n[d] + 2) if d in wd else slice(None) for d in range(3))
    return np.asfortranarray(out[isl])


def scal_jt(case, qh, sh, sigma, alpha, cbc):
    """Transposed Jacobians of the scalar right-hand side of scal.f90:41-98 (explicit diffusion),
        S(x) = -sum_d H_d(k) [F_d(x) - F_d(x - e_d)] + alpha L s,  F_d(x) = (s(x + e_d) + s(x))/2 q_d(x),
    H = dxi, dyi, dzfi(k): with R_d(x) = H_d(k_x) sigma(x) - H_d(k_x+e_d) sigma(x + e_d),
        (J_s^T sigma)(y) = -1/2 sum_d [q_d(y - e_d) R_d(y - e_d) + q_d(y) R_d(y)] + alpha L^T sigma(y),
        (J_q^T sigma)_d(y) = -1/2 (s(y + e_d) + s(y)) R_d(y).
    qh, sh: the halo'd velocity and scalar (ghosts as the forward set them); sigma: interior array.
    Returns (J_s^T sigma folded by the scalar's ghost map, [J_q^T sigma] on the velocity's degrees of
    freedom)."""
    Q = [_ext(case, x, True) for x in qh]
    S = _ext(case, sh, True)
    G = _ext(case, sigma, False)
    zc_, zf_ = _z_tables(case)
    H = [case.dli[0], case.dli[1], zf_]
    out = np.zeros_like(G)
    qt = []
    for d in range(3):
        hs = H[d] * G
        R = hs - _sh(hs, d, 1)
        out -= 0.5 * (_sh(Q[d] * R, d, -1) + Q[d] * R)
        qt.append(-0.5 * (_sh(S, d, 1) + S) * R)
    xy = sum(case.dli[d] ** 2 * (_sh(G, d, 1) - 2 * G + _sh(G, d, -1)) for d in range(2))
    gh = zf_ * G
    z = zc_ * (_sh(gh, 2, 1) - gh) - _sh(zc_, 2, -1) * (gh - _sh(gh, 2, -1))
    out += alpha * (xy + z)
    n, wd = case.n, wall_dirs(case)
    isl = tuple(slice(2, n[d] + 2) if d in wd else slice(None) for d in range(3))
    qres = [np.asfortranarray(x[isl]) for x in qt]
    _clear_wall_faces(case, qres)
    return _fold_scalar(case, out, cbc), qres


def buoyancy_t(case, wb, cbc, coef):
    """Transpose of the Boussinesq term u_c -= coef_c (s(x + e_c) + s(x)) (rk.f90:163-177, coef_c =
    f12 gacc_c beta / 2) applied to wb, folded by the scalar's ghost map: sum_c -coef_c (wb_c(y - e_c) + wb_c(y)),
    wb zero off the velocity's degrees of freedom."""
    wz = [x.copy() for x in wb]
    _clear_wall_faces(case, wz)
    W = [_ext(case, x, False) for x in wz]
    out = np.zeros_like(W[0])
    for c in range(3):
        out -= coef[c] * (_sh(W[c], c, -1) + W[c])
    return _fold_scalar(case, out, cbc)


def kinetic_energy(case, u, v, w):
    """Volume-averaged kinetic energy of the interior degrees of freedom (dzf- and dzc-weighted)."""
    wts = volume_weights(case)
    vol = float(np.sum(case.dzf[1:case.n[2] + 1])) * case.n[0] * case.n[1]
    return 0.5 * sum(float(np.sum(m * q * q)) for m, q in zip(wts, (u, v, w))) / vol


def ke_grad(case, u, v, w):
    """d kinetic_energy / d(u, v, w)."""
    wts = volume_weights(case)
    vol = float(np.sum(case.dzf[1:case.n[2] + 1])) * case.n[0] * case.n[1]
    return [np.asfortranarray(m * q / vol) for m, q in zip(wts, (u, v, w))]


def _state(case, u0, v0, w0, scal0=()):
    st = case.new_state()
    for k, q in zip("uvw", (u0, v0, w0)):
        st[k][1:-1, 1:-1, 1:-1] = q
    for i, q in enumerate(scal0):
        st["s%d" % i][1:-1, 1:-1, 1:-1] = q
    case.init_bounds(st)
    return st


def _step_recording(case, st, dt):
    """One RK3 step (the oracle's rk3_step sequence) returning, for each substep, the halo'd velocity and
    scalars at its start."""
    u, v, w, p, pp = st["u"], st["v"], st["w"], st["p"], st["pp"]
    starts = []
    for irk in range(3):
        starts.append(([x.copy(order="F") for x in (u, v, w)],
                       [st["s%d" % i].copy(order="F") for i in range(len(case.scalars))]))
        rkpar = O.RKCOEFF[irk]
        dtrk = (rkpar[0] + rkpar[1]) * dt
        for i, sc in enumerate(case.scalars):  # main.f90:440-450
            sf = st["s%d" % i]
            fs = case.rk_scal(rkpar, dt, sc, u, v, w, sf, st["dsdtrko%d" % i])
            case.bulk_forcing_s(sc["is_forced"], fs, sf)
            case.boundp(sf, cbc=sc["cbc"], bc=sc["bc"])
        case.rk(rkpar, dt, p, st["dudtrko"], st["dvdtrko"], st["dwdtrko"], u, v, w, st.get("s0"))
        alpha = 0.0
        if case.is_impdiff:  # main.f90:453-462
            alpha = -.5 * case.visc * dtrk
            for q, h in zip((u, v, w), case.h_vel):
                case.solve_helmholtz(q, alpha, h)
        case.bounduvw(u, v, w, 0, -1)
        case.fillps(u, v, w, pp, dtrk ** (-1))
        case.updt_rhs_b(pp)
        case.solver(pp)
        case.boundp(pp)
        case.correc(pp, u, v, w, dtrk)
        case.bounduvw(u, v, w, 1, 1, 0)
        case.updatep(pp, p, alpha)
        case.boundp(p, 1)
    return starts


def helmholtz_inv(case, a, dtrk):
    """H^-1 a on each component, H = I - visc dtrk/2 L the implicit-diffusion operator of the substep
    (main.f90:453-462; symmetric on a tri-periodic uniform grid, so H^-T = H^-1)."""
    out = []
    for c, x in enumerate(a):
        f = halo(case, x)
        case.solve_helmholtz(f, -0.5 * case.visc * dtrk, case.h_vel[c])
        out.append(interior(f))
    return out


def substep_adjoint(case, start, irk, dt, ub, fb, sb=(), fsb=()):
    """Adjoint of substep irk (0-based): start = the halo'd velocity (and, as (velocity, scalars), the
    scalars) it started from; (ub, fb[, sb, fsb]) -> (ub, fb[, sb, fsb]), the adjoints of the velocity, its
    stored right-hand side and, per scalar, of the scalar and its stored right-hand side.

    With implicit diffusion (3-D, tri-periodic uniform grid) the substep is
    u' = P H^-1 Fr(u + f1 A(u) + f2 Ao + (f1 + f2)/2 visc L u), Ao' = A(u), A the advection alone
    (rk.f90:156-186, 312-326), so its adjoint is hb = Fr'^T H^-1 P ub,
    ub <- hb + J_A^T (f1 hb + fb) + (f1 + f2)/2 visc L hb, fb <- f2 hb.

    Scalars (explicit diffusion; main.f90:440-450): each is advanced first,
    s' = Fs(s + f1 S(u, s) + f2 So), So' = S(u, s), Fs its bulk forcing when forced, and the velocity's rk
    then reads the first one's new value in the Boussinesq term -f12 gacc beta (s(x + e_c) + s(x))/2
    (rk.f90:163-177).  So the first scalar's adjoint gains that term's transpose of wb before the
    scalars' own adjoints sh = Fs'^T (sb (+ buoyancy^T wb)), sb <- sh + J_s^T (f1 sh + fsb),
    fsb <- f2 sh, ub += J_u^T (f1 sh + fsb)."""
    vel, scal = start if isinstance(start, tuple) else (start, [])
    f1, f2 = O.RKCOEFF[irk][0] * dt, O.RKCOEFF[irk][1] * dt
    if case.is_impdiff:
        hb = forcing_t(case, helmholtz_inv(case, project_t(case, ub), f1 + f2))
        g = [f1 * a + b for a, b in zip(hb, fb)]
        jt = mom_jt(case, [interior(x) for x in vel], g, visc=0.)
        half = 0.5 * (f1 + f2) * case.visc
        return [a + b + half * _lap(case, a) for a, b in zip(hb, jt)], [f2 * a for a in hb]
    wb = forcing_t(case, project_t(case, ub))
    g = [f1 * a + b for a, b in zip(wb, fb)]
    if walls(case):
        jt = mom_jt_walls(case, vel, g)
    else:
        jt = mom_jt(case, [interior(x) for x in vel], g)
    ub_new, fb_new = [a + b for a, b in zip(wb, jt)], [f2 * a for a in wb]
    if not case.scalars:
        return ub_new, fb_new
    sh = [np.asarray(x, dtype=np.float64).copy() for x in sb]
    if case.is_buoyancy:
        coef = [(f1 + f2) * case.gacc[c] * case.beta * 0.5 for c in range(3)]
        sh[0] = sh[0] + buoyancy_t(case, wb, case.scalars[0]["cbc"], coef)
    gvr = np.asarray(case.grid_vol_ratio_f[1:case.n[2] + 1], dtype=np.float64)[None, None, :]
    for i, sc in enumerate(case.scalars):
        if sc["is_forced"]:  # s <- s + (scalf - sum gvr_f s) (rk.f90:474-475, scal.f90:313-329): I - 1 gvr_f^T
            sh[i] = sh[i] - gvr * float(np.sum(sh[i]))
    sb_new, fsb_new = [], []
    for i, sc in enumerate(case.scalars):
        sigma = f1 * sh[i] + fsb[i]
        js, jq = scal_jt(case, vel, scal[i], sigma, sc["alpha"], sc["cbc"])
        sb_new.append(sh[i] + js)
        fsb_new.append(f2 * sh[i])
        ub_new = [a + b for a, b in zip(ub_new, jq)]
    return ub_new, fb_new, sb_new, fsb_new


def step_adjoint(case, starts, ub, dt):
    """Adjoint of one step: ub (adjoint of the step's final velocity) -> adjoint of its initial velocity
    (the scalars' adjoints start at zero: the loss reads the velocity only)."""
    fb = [np.zeros_like(x) for x in ub]
    ns = len(case.scalars)
    sb = [np.zeros_like(ub[0]) for _ in range(ns)]
    fsb = [np.zeros_like(ub[0]) for _ in range(ns)]
    for irk in (2, 1, 0):
        r = substep_adjoint(case, starts[irk], irk, dt, ub, fb, sb, fsb)
        ub, fb = r[0], r[1]
        if ns:
            sb, fsb = r[2], r[3]
    return ub, sb


def ke_after_steps(case, u0, v0, w0, dt, nsteps, scal0=()):
    """Kinetic energy after nsteps RK3 steps from the interior velocity (u0, v0, w0) (and scalars scal0),
    zero pressure."""
    st = _state(case, u0, v0, w0, scal0)
    for _ in range(nsteps):
        case.rk3_step(st, dt)
    return kinetic_energy(case, *(interior(st[k]) for k in "uvw"))


def grad_ke_after_steps(case, u0, v0, w0, dt, nsteps, every=4, scal0=()):
    """(E, dE/du0): E the kinetic energy after nsteps steps; checkpoints of the state every `every`
    steps, the steps in between recomputed during the backward sweep.  With scalars the adjoint carries
    their adjoints across steps too."""
    check_case(case)
    st = _state(case, u0, v0, w0, scal0)
    ckpt = {}
    for s in range(nsteps):
        if s % every == 0:
            ckpt[s] = {k: v.copy(order="F") for k, v in st.items()}
        case.rk3_step(st, dt)
    fin = [interior(st[k]) for k in "uvw"]
    E = kinetic_energy(case, *fin)
    ub = ke_grad(case, *fin)
    ns = len(case.scalars)
    sb = [np.zeros_like(ub[0]) for _ in range(ns)]
    for s0 in sorted(ckpt, reverse=True):
        seg = {k: v.copy(order="F") for k, v in ckpt[s0].items()}
        rec = [_step_recording(case, seg, dt) for _ in range(s0, min(s0 + every, nsteps))]
        for starts in reversed(rec):
            fb = [np.zeros_like(x) for x in ub]
            fsb = [np.zeros_like(ub[0]) for _ in range(ns)]
            for irk in (2, 1, 0):
                r = substep_adjoint(case, starts[irk], irk, dt, ub, fb, sb, fsb)
                ub, fb = r[0], r[1]
                if ns:
                    sb, fsb = r[2], r[3]
    return E, ub
