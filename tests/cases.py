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
    elif name == "cavity_dd":  # Dirichlet pressure walls in y (DST), Neumann in z
        p = dict(ng=(32, 32, 24), l=(1.0, 1.0, 1.0), gtype=1, gr=0.5, cbcvel="PPNNDD" * 3,
                 cbcpre="PPDDNN", visci=100.0)
    else:
        raise KeyError(name)
    if ng is not None:
        p["ng"] = tuple(ng)
    return p


def oracle_case(p):
    from oracle import oracle as O
    return O.OracleCase(ng=p["ng"], l=p["l"], gtype=p["gtype"], gr=p["gr"], cbcvel=p["cbcvel"],
                        cbcpre=p["cbcpre"], bcvel=p.get("bcvel"), visci=p["visci"],
                        is_forced=p.get("is_forced", (0, 0, 0)), velf=p.get("velf", (0, 0, 0)))


def host_case(p):
    from cans_amd.host import Case
    return Case(ng=p["ng"], l=p["l"], gtype=p["gtype"], gr=p["gr"], cbcvel=p["cbcvel"], cbcpre=p["cbcpre"],
                bcvel=tuple(p.get("bcvel", (0.0,) * 18)), visci=p["visci"],
                is_forced=tuple(bool(x) for x in p.get("is_forced", (0, 0, 0))),
                velf=tuple(p.get("velf", (0.0, 0.0, 0.0))))


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
