"""Case definitions shared by the parity tests and bench.py.

Each case is given once as plain parameters and turned into both the oracle's
OracleCase and the device host's Case, so both sides see identical inputs.
"""
import numpy as np

from cans_amd import initflow as IF


def params(name, ng=None):
    """Named workloads.  `ng` overrides the grid (parity tests use small grids)."""
    if name == "tgv":  # tri-periodic Taylor-Green box of side 2*pi, nu = 1/1600
        p = dict(ng=(128, 128, 128), l=(2 * np.pi,) * 3, gtype=1, gr=0.0, cbcvel="PP" * 9,
                 cbcpre="PPPPPP", visci=1600.0)
    elif name == "channel":  # periodic x,y; no-slip walls in z; tanh-stretched z; constant flow rate
        p = dict(ng=(512, 256, 192), l=(4 * np.pi, 2 * np.pi, 1.0), gtype=1, gr=1.8,
                 cbcvel="PPPPDD" * 3, cbcpre="PPPPNN", visci=5600.0, is_forced=(1, 0, 0),
                 velf=(1.0, 0.0, 0.0))
    elif name == "ldc":  # the reference's lid-driven cavity regression case
        bcvel = np.zeros(18)
        bcvel[1 + 4 + 6] = 1.0
        p = dict(ng=(2, 64, 64), l=(0.03125, 1.0, 1.0), gtype=1, gr=0.0, cbcvel="PPDDDD" * 3,
                 cbcpre="PPNNNN", bcvel=tuple(bcvel), visci=1000.0)
    elif name == "duct":  # periodic x, walls in y and z
        p = dict(ng=(64, 32, 48), l=(2.0, 1.0, 1.0), gtype=1, gr=1.2, cbcvel="PPDDDD" * 3,
                 cbcpre="PPNNNN", visci=400.0, is_forced=(1, 0, 0), velf=(1.0, 0.0, 0.0))
    elif name == "box":  # no-slip walls on all six faces (Neumann pressure everywhere, DCT in x and y)
        p = dict(ng=(16, 16, 12), l=(1.0, 1.0, 1.0), gtype=1, gr=0.5, cbcvel="DDDDDD" * 3,
                 cbcpre="NNNNNN", visci=100.0)
    elif name == "open_nd":  # outflow faces at x = lx and y = 0 (velocity N, pressure D): pressure ND in x
        # (REDFT11), DN in y (RODFT11), walls in z
        p = dict(ng=(16, 24, 12), l=(1.0, 1.0, 1.0), gtype=1, gr=0.5, cbcvel="DNNDDD" * 3,
                 cbcpre="NDDNNN", visci=100.0)
    elif name == "open_x":  # outflow faces at both x ends (velocity N, pressure D: DST-II along x), walls in z
        p = dict(ng=(64, 16, 12), l=(1.0, 1.0, 1.0), gtype=1, gr=0.5, cbcvel="NNPPDD" * 3,
                 cbcpre="DDPPNN", visci=100.0)
    elif name == "cavity_dd":  # Dirichlet pressure walls in y (DST), Neumann in z
        p = dict(ng=(32, 32, 24), l=(1.0, 1.0, 1.0), gtype=1, gr=0.5, cbcvel="PPNNDD" * 3,
                 cbcpre="PPDDNN", visci=100.0)
    elif name == "dhc":  # the reference's differentially heated cavity (tests/differentially_heated_cavity)
        p = dict(ng=(128, 2, 128), l=(1.0, 0.015625, 1.0), gtype=1, gr=0.0, cbcvel="DDPPDD" * 3,
                 cbcpre="NNPPNN", visci=1186.78165819,
                 scalars=(dict(alphai=842.614977318, cbc="DDPPNN", bc=(-0.5, 0.5, 0.0, 0.0, 0.0, 0.0)),),
                 gacc=(0.0, 0.0, -1.0), beta=1.0, is_boussinesq_buoyancy=True)
    elif name == "dhc3d":  # heated cube: no-slip walls, hot/cold x walls, adiabatic y/z walls, implicit z diffusion
        p = dict(ng=(256, 256, 256), l=(1.0, 1.0, 1.0), gtype=1, gr=0.0, cbcvel="DDDDDD" * 3,
                 cbcpre="NNNNNN", visci=1186.78165819,
                 scalars=(dict(alphai=842.614977318, cbc="DDNNNN", bc=(-0.5, 0.5, 0.0, 0.0, 0.0, 0.0)),),
                 gacc=(0.0, 0.0, -1.0), beta=1.0, is_boussinesq_buoyancy=True, is_impdiff_1d=True)
    elif name == "rb_channel":  # heated channel: scalar walls in z (Dirichlet), buoyancy, two scalars, one forced
        p = dict(ng=(32, 16, 24), l=(2.0, 1.0, 1.0), gtype=1, gr=1.2, cbcvel="PPPPDD" * 3,
                 cbcpre="PPPPNN", visci=300.0, is_forced=(1, 0, 0), velf=(1.0, 0.0, 0.0),
                 scalars=(dict(alphai=250.0, cbc="PPPPDD", bc=(0.0, 0.0, 0.0, 0.0, 1.0, -1.0), source=0.3),
                          dict(alphai=400.0, cbc="PPPPNN", bc=(0.0, 0.0, 0.0, 0.0, 0.5, -0.25), is_forced=True,
                               scalf=0.7)),
                 gacc=(0.1, 0.0, -2.0), beta=0.8, is_boussinesq_buoyancy=True)
    else:
        raise KeyError(name)
    if ng is not None:
        p["ng"] = tuple(ng)
    return p


def oracle_case(p):
    from oracle import oracle as O
    return O.OracleCase(ng=p["ng"], l=p["l"], gtype=p["gtype"], gr=p["gr"], cbcvel=p["cbcvel"],
                        cbcpre=p["cbcpre"], bcvel=p.get("bcvel"), visci=p["visci"],
                        is_forced=p.get("is_forced", (0, 0, 0)), velf=p.get("velf", (0, 0, 0)),
                        scalars=p.get("scalars", ()), gacc=p.get("gacc", (0, 0, 0)), beta=p.get("beta", 0.0),
                        is_boussinesq_buoyancy=p.get("is_boussinesq_buoyancy", False),
                        is_impdiff=p.get("is_impdiff", False) or p.get("is_impdiff_1d", False),
                        is_impdiff_1d=p.get("is_impdiff_1d", False),
                        is_gridpoint_natural_channel=bool(p.get("natural_grid", False)))


def host_case(p):
    from cans_amd.host import Case
    return Case(ng=p["ng"], l=p["l"], gtype=p["gtype"], gr=p["gr"], cbcvel=p["cbcvel"], cbcpre=p["cbcpre"],
                bcvel=tuple(p.get("bcvel", (0.0,) * 18)), visci=p["visci"],
                is_forced=tuple(bool(x) for x in p.get("is_forced", (0, 0, 0))),
                velf=tuple(p.get("velf", (0.0, 0.0, 0.0))), scalars=tuple(p.get("scalars", ())),
                gacc=tuple(p.get("gacc", (0.0, 0.0, 0.0))), beta=float(p.get("beta", 0.0)),
                is_boussinesq_buoyancy=bool(p.get("is_boussinesq_buoyancy", False)),
                is_impdiff=bool(p.get("is_impdiff", False) or p.get("is_impdiff_1d", False)),
                is_impdiff_1d=bool(p.get("is_impdiff_1d", False)), dtdma_slabs=int(p.get("dtdma_slabs", 0)),
                is_gridpoint_natural_channel=bool(p.get("natural_grid", False)))


def initial_state(name, p, grid, seed=2024):
    """Interior (u, v, w, p) for a case, seeded (see cans_amd.initflow)."""
    n = p["ng"]
    g = IF.rng(seed)
    if name == "tgv":
        u, v, w = IF.tgv(n, grid.dl, grid.l, grid.zc)
        for q in (u, v, w):
            IF.add_noise(q, g, 1e-2)
    elif name in ("channel", "duct"):
        u1d = IF.poiseuille(grid.zc[1:n[2] + 1] / grid.l[2], 1.0)
        u = np.asfortranarray(np.broadcast_to(u1d, tuple(n)).copy())
        v = np.zeros(tuple(n), order="F")
        w = np.zeros(tuple(n), order="F")
        for q in (u, v, w):
            IF.add_noise(q, g, 0.1)
        gvr = grid.dzf[1:n[2] + 1] / grid.l[2] * (grid.dl[0] / grid.l[0]) * (grid.dl[1] / grid.l[1])
        IF.set_mean(u, gvr, 1.0)
    else:
        u, v, w = (np.asfortranarray(g.uniform(-0.1, 0.1, tuple(n))) for _ in range(3))
    pr = np.zeros(tuple(n), order="F")
    return u, v, w, pr


def initial_scalars(name, p, grid, seed=2025):
    """Interior fields of the case's scalars.  dhc, dhc3d: the linear profile between the
    x walls (initflow.f90:327-337); otherwise seeded noise."""
    n = p["ng"]
    out = []
    g = IF.rng(seed)
    for sc in p.get("scalars", ()):
        if name in ("dhc", "dhc3d"):
            xx = (np.arange(1, n[0] + 1) - 0.5) * grid.dl[0] / grid.l[0]
            prof = (1.0 - xx) * sc["bc"][0] + xx * sc["bc"][1]
            out.append(np.asfortranarray(np.broadcast_to(prof[:, None, None], tuple(n)).copy()))
        else:
            out.append(np.asfortranarray(g.uniform(-0.5, 0.5, tuple(n))))
    return out


def dhc_nusselt(s_interior, dl):
    """Nusselt number of tests/differentially_heated_cavity/test.py:7-14 from s(1,1,:):
    tw = -0.5, l = 1, dx = 2*xp(1), dz = 2*zp(1) on the uniform grid."""
    tw, l = -0.5, 1.0
    dx, dz = dl[0], dl[2]
    return float((1.0 / l) * np.sum(((s_interior[0, 0, :] - tw) / dx) * dz) * l / (-tw))


def poisson_boxes(nbox, seed, sizes=(32, 48, 64, 96, 128), nz=(32, 160), gr=(0.0, 2.0), cbcs=("PPPPNN",)):
    """Seeded independent Poisson boxes: nx, ny drawn from `sizes`, nz uniform in `nz`
    (inclusive), tanh stretching factor uniform in `gr`, pressure BCs drawn from `cbcs`.
    Each box spans (2*pi, 2*pi, 1) like a channel cell."""
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(nbox):
        n1, n2 = (int(rng.choice(sizes)) for _ in range(2))
        n3 = int(rng.integers(nz[0], nz[1] + 1))
        out.append(dict(n=(n1, n2, n3), l=(2 * np.pi, 2 * np.pi, 1.0), gtype=1, gr=float(rng.uniform(*gr)),
                        cbc=cbcs[int(rng.integers(len(cbcs)))]))
    return out


def box_rhs(box, dzf, rng):
    """uniform[-1, 1] rhs of a box with its volume-weighted mean removed when no wall is Dirichlet."""
    rhs = rng.uniform(-1, 1, tuple(box["n"]))
    if "D" not in box["cbc"]:
        wgt = np.broadcast_to(dzf, rhs.shape)
        rhs -= (rhs * wgt).sum() / wgt.sum()
    return rhs
