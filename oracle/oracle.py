"""ctypes front-end of the C oracle (TEST INFRASTRUCTURE ONLY).

The oracle is the checker for the HIP path and the timed CPU baseline.  Only
``tests/``, ``__graft_entry__.smoke()`` and ``bench.py``'s cpu_baseline leg may
import this module.  It never imports the product package, so that a mistake
in the product's host setup cannot hide behind a shared helper.

``OracleCase`` mirrors the reference's initialisation (src/main.f90:241-297)
and its RK3 substep sequence (src/main.f90:437-473) on top of the C routines.
Fields are numpy arrays of shape (n1+2, n2+2, n3+2) in Fortran order, i.e.
the reference's halo'd ``u(0:n1+1,0:n2+1,0:n3+1)``.
"""
import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIBPATH = os.path.join(_HERE, "liboracle.so")

R2HC, HC2R, DHT, REDFT00, REDFT01, REDFT10, REDFT11, RODFT00, RODFT01, RODFT10, RODFT11 = range(11)
KIND_NAMES = {R2HC: "R2HC", HC2R: "HC2R", REDFT00: "REDFT00", REDFT01: "REDFT01",
              REDFT10: "REDFT10", REDFT11: "REDFT11", RODFT00: "RODFT00",
              RODFT01: "RODFT01", RODFT10: "RODFT10", RODFT11: "RODFT11"}

# rk coefficients, src/param.f90:20-23
RKCOEFF = np.array([[32. / 60., 0.], [25. / 60., -17. / 60.], [45. / 60., -25. / 60.]])

_lib = None


def build():
    subprocess.check_call(["make", "-s", "-C", _HERE])


def use_native():
    """Build the oracle with -march=native for this host and load that build (the CPU
    baseline's build; call before the first lib()).  Returns the compiler flags' -march."""
    global _LIBPATH
    if _lib is not None:
        raise RuntimeError("use_native() must come before the oracle is loaded")
    subprocess.check_call(["make", "-s", "-B", "-C", _HERE, "liboracle_native.so"])
    _LIBPATH = os.path.join(_HERE, "liboracle_native.so")
    return "native"


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIBPATH):
            build()
        _lib = C.CDLL(_LIBPATH)
        _declare(_lib)
    return _lib


_dp = C.POINTER(C.c_double)
_ip = C.POINTER(C.c_int)


def _declare(L):
    L.orc_r2r.argtypes = [C.c_int, C.c_int, C.c_long, C.c_long, C.c_long, _dp]
    L.orc_initgrid.argtypes = [C.c_int, C.c_int, C.c_double, C.c_double, C.c_int, _dp, _dp, _dp, _dp]
    L.orc_initgrid_natural.argtypes = [C.c_int, C.c_double, C.c_int, _dp, _dp, _dp, _dp]
    L.orc_eigenvalues.argtypes = [C.c_int, C.c_char_p, C.c_char, _dp]
    L.orc_tridmatrix.argtypes = [C.c_char_p, C.c_int, _dp, _dp, C.c_char, _dp, _dp, _dp]
    L.orc_bc_rhs.argtypes = [C.c_char_p, _dp, _dp, _dp, C.c_char, _dp]
    L.orc_fftini.argtypes = [_ip, C.c_char_p, C.c_char_p, _ip, _ip, _dp]
    L.orc_gaussel.argtypes = [C.c_int, C.c_int, C.c_int, _dp, _dp, _dp, C.c_int, C.c_double, _dp, C.c_void_p]
    L.orc_solver.argtypes = [_ip, _ip, _ip, C.c_double, _dp, _dp, _dp, _dp, C.c_char_p, C.c_char_p, _dp]
    L.orc_solve_helmholtz.argtypes = [_ip, _ip, _ip, C.c_double, C.c_double, _dp, _dp, _dp, _dp,
                                      C.c_void_p, C.c_void_p, C.c_void_p, _ip, C.c_char_p, C.c_char_p, _dp]
    L.orc_mom_xyz_ad.argtypes = [_ip, C.c_double, C.c_double, _dp, _dp, C.c_double, _dp, _dp, _dp,
                                 _dp, _dp, _dp, C.c_void_p, C.c_void_p, C.c_void_p, C.c_int, C.c_int]
    L.orc_rk.argtypes = [_dp, _ip, _dp, _dp, _dp, _dp, _dp, C.c_double, C.c_double, _dp, _ip, _dp, _dp,
                         _dp, C.c_double, C.c_void_p,
                         _dp, _dp, _dp, _dp, _dp, _dp, _dp, _dp, C.c_int, C.c_int]
    L.orc_scal.argtypes = [_ip, C.c_double, C.c_double, _dp, _dp, C.c_double, _dp, _dp, _dp, _dp, _dp,
                           C.c_void_p, C.c_int, C.c_int]
    L.orc_rk_scal.argtypes = [_dp, _ip, _dp, _dp, _dp, _dp, C.c_double, C.c_double, _dp, _dp, _dp, C.c_int,
                              C.c_double, C.c_double, _dp, _dp, _dp, _dp, C.c_int, C.c_int]
    L.orc_bulk_forcing_s.argtypes = [_ip, C.c_int, C.c_double, _dp]
    L.orc_gaussel_dtdma.argtypes = [C.c_int, C.c_int, C.c_int, C.c_int, _dp, _dp, _dp, C.c_double, _dp,
                                    C.c_void_p, C.c_int]
    L.orc_solver_gaussel_z.argtypes = [_ip, _dp, _dp, _dp, C.c_char_p, C.c_char, C.c_double, _dp]
    L.orc_solve_helmholtz_1d.argtypes = [_ip, C.c_double, _dp, _dp, _dp, C.c_void_p, _ip, C.c_char_p,
                                         C.c_char_p, _dp]
    L.orc_bulk_mean.argtypes = [_ip, _dp, _dp]
    L.orc_bulk_mean.restype = C.c_double
    L.orc_fillps.argtypes = [_ip, _dp, _dp, C.c_double, _dp, _dp, _dp, _dp]
    L.orc_correc.argtypes = [_ip, _dp, _dp, C.c_double, _dp, _dp, _dp, _dp]
    L.orc_updatep.argtypes = [_ip, _dp, _dp, _dp, C.c_double, _dp, _dp, C.c_int, C.c_int]
    L.orc_bounduvw.argtypes = [C.c_char_p, _ip, _dp, _ip, _ip, _dp, _dp, _dp, _dp, _dp, _dp,
                               C.c_int, C.c_int, C.c_int]
    L.orc_boundp.argtypes = [C.c_char_p, _ip, _dp, _ip, _ip, _dp, _dp, _dp, C.c_int]
    L.orc_updt_rhs_b.argtypes = [C.c_char_p, C.c_char_p, _ip, _ip, C.c_void_p, C.c_void_p, C.c_void_p,
                                 _dp, C.c_double, C.c_int]
    L.orc_chkdiv.argtypes = [_ip, _dp, _dp, _dp, _dp, _dp, _dp, _dp, _dp]
    L.orc_chkdt.argtypes = [_ip, _dp, _dp, _dp, C.c_double, C.c_double, _dp, _dp, _dp, C.c_int, C.c_int]
    L.orc_chkdt.restype = C.c_double
    L.orc_num_threads.restype = C.c_int
    L.orc_set_fft_variant.argtypes = [C.c_int]
    L.orc_set_threads.argtypes = [C.c_int]


def fft_variant(v):
    lib().orc_set_fft_variant(int(v))


def set_threads(n):
    lib().orc_set_threads(int(n))


def P(a):
    """double* of a contiguous float64 numpy array (Fortran- or C-contiguous)."""
    assert a.dtype == np.float64 and (a.flags.f_contiguous or a.flags.c_contiguous)
    return a.ctypes.data_as(_dp)


def IP(seq):
    arr = (C.c_int * len(seq))(*[int(x) for x in seq])
    return C.cast(arr, _ip), arr


def DP(seq):
    arr = (C.c_double * len(seq))(*[float(x) for x in seq])
    return C.cast(arr, _dp), arr


def VP(a):
    return None if a is None else C.c_void_p(a.ctypes.data)


def chars(s):
    return s.encode("ascii") if isinstance(s, str) else bytes(s)


# ---------------------------------------------------------------------------
# thin wrappers
# ---------------------------------------------------------------------------
def r2r(kind, x, axis=-1):
    """FFTW-definition r2r transform of a float64 array along ``axis`` (copy)."""
    y = np.ascontiguousarray(np.moveaxis(np.array(x, dtype=np.float64), axis, -1))
    n = y.shape[-1]
    howmany = y.size // n if n else 0
    rc = lib().orc_r2r(kind, n, howmany, 1, n, P(y))
    if rc != 0:
        raise ValueError("orc_r2r failed (%d)" % rc)
    return np.moveaxis(y, -1, axis)


def initgrid(gtype, n, gr, lz, is_periodic):
    dzc, dzf, zc, zf = (np.zeros(n + 2) for _ in range(4))
    lib().orc_initgrid(gtype, n, gr, lz, int(is_periodic), P(dzc), P(dzf), P(zc), P(zf))
    return dzc, dzf, zc, zf


def initgrid_natural(n, lz, is_periodic):
    dzc, dzf, zc, zf = (np.zeros(n + 2) for _ in range(4))
    lib().orc_initgrid_natural(n, lz, int(is_periodic), P(dzc), P(dzf), P(zc), P(zf))
    return dzc, dzf, zc, zf


def eigenvalues(n, bc, c_or_f):
    lam = np.zeros(n)
    lib().orc_eigenvalues(n, chars(bc), chars(c_or_f), P(lam))
    return lam


def tridmatrix(bc, n, dzci, dzfi, c_or_f):
    a, b, c = np.zeros(n), np.zeros(n), np.zeros(n)
    lib().orc_tridmatrix(chars(bc), n, P(dzci), P(dzfi), chars(c_or_f), P(a), P(b), P(c))
    return a, b, c


def gaussel(p, a, b, c, is_periodic, norm, lambdaxy=None, n=None):
    """p: (nx, ny, nz) Fortran-order array, solved in place on its first n planes."""
    nx, ny, nz = p.shape
    assert p.flags.f_contiguous
    lam = None if lambdaxy is None else np.asfortranarray(lambdaxy, dtype=np.float64)
    lib().orc_gaussel(nx, ny, nz if n is None else n, P(a), P(b), P(c), int(is_periodic), norm, P(p),
                      VP(lam))
    return p


def fftini(ng, bcxy, c_or_f):
    kinds = (C.c_int * 4)()
    lens = (C.c_int * 2)()
    nf = C.c_double()
    ngp, _k = IP(ng)
    rc = lib().orc_fftini(ngp, chars(bcxy), chars(c_or_f), kinds, lens, C.byref(nf))
    if rc:
        raise ValueError("unsupported boundary conditions %r" % (bcxy,))
    return list(kinds), list(lens), nf.value


def field(n, fill=0.0):
    return np.full((n[0] + 2, n[1] + 2, n[2] + 2), fill, dtype=np.float64, order="F")


# ---------------------------------------------------------------------------
# a case: reference initialisation + time step on top of the C routines
# ---------------------------------------------------------------------------
class OracleCase:
    """State-free description of one CaNS case on a single rank.

    cbcvel: 18 chars flattening cbcvel(0:1,3,3); cbcpre: 6 chars; bcvel(18)/bcpre(6) floats.
    """

    def __init__(self, ng, l, gtype=1, gr=0.0, cbcvel="PP" * 9, cbcpre="PP" * 3,
                 bcvel=None, bcpre=None, visci=1.0, bforce=(0, 0, 0), is_forced=(0, 0, 0),
                 velf=(0, 0, 0), is_impdiff=False, is_impdiff_1d=False, scalars=(), gacc=(0, 0, 0),
                 beta=0.0, is_boussinesq_buoyancy=False, is_gridpoint_natural_channel=False):
        L = lib()
        self.ng = [int(x) for x in ng]
        self.n = list(self.ng)
        self.l = [float(x) for x in l]
        # param.f90:196-198 -- (1.*ng) is default real
        self.dl = [self.l[d] / float(np.float32(self.ng[d])) for d in range(3)]
        self.dli = [self.dl[d] ** (-1) for d in range(3)]
        self.visc = float(visci) ** (-1)
        self.cbcvel = cbcvel
        self.cbcpre = cbcpre
        self.bcvel = np.zeros(18) if bcvel is None else np.asarray(bcvel, dtype=np.float64)
        self.bcpre = np.zeros(6) if bcpre is None else np.asarray(bcpre, dtype=np.float64)
        self.bforce = [float(x) for x in bforce]
        self.is_forced = [int(bool(x)) for x in is_forced]
        self.velf = [float(x) for x in velf]
        self.is_impdiff = int(is_impdiff)
        self.is_impdiff_1d = int(is_impdiff_1d)
        self.is_periodic = [int(cbcpre[2 * d] == "P" and cbcpre[2 * d + 1] == "P") for d in range(3)]
        # single rank, x pencils: is_bound along x always, elsewhere iff not periodic (initmpi.f90:264-267)
        self.is_bound = [1, 1] + [0 if self.is_periodic[d] else 1 for d in (1, 1, 2, 2)]
        n3 = self.n[2]
        # main.f90:241-285
        if is_gridpoint_natural_channel:
            self.dzc, self.dzf, self.zc, self.zf = initgrid_natural(n3, self.l[2], cbcpre[4:6] == "PP")
        else:
            self.dzc, self.dzf, self.zc, self.zf = initgrid(gtype, n3, gr, self.l[2], cbcpre[4:6] == "PP")
        self.dzci = self.dzc ** (-1)
        self.dzfi = self.dzf ** (-1)
        vol = self.l[0] * self.l[1] * self.l[2]
        self.grid_vol_ratio_c = self.dl[0] * self.dl[1] * self.dzc / vol
        self.grid_vol_ratio_f = self.dl[0] * self.dl[1] * self.dzf / vol
        self.p_solver = self.initsolver(cbcpre, self.bcpre, "ccc")
        self.work = np.zeros(3 * int(np.prod(self.n)) * (2 if is_impdiff else 1))
        # scalars (scal.f90:277-311): alphai, cbc (6 chars), bc (6), source, is_forced, scalf
        self.scalars = []
        for sc in scalars:
            self.scalars.append(dict(alpha=float(sc.get("alphai", 1.0)) ** (-1), cbc=sc["cbc"],
                                     bc=np.asarray(sc.get("bc", np.zeros(6)), dtype=np.float64),
                                     source=float(sc.get("source", 0.0)), is_forced=int(bool(sc.get("is_forced", 0))),
                                     scalf=float(sc.get("scalf", 0.0))))
        self.gacc = [float(x) for x in gacc]
        self.beta = float(beta)
        self.is_buoyancy = bool(is_boussinesq_buoyancy) and len(self.scalars) > 0
        # param.f90:308-310: the largest scalar diffusivity (1/huge without scalars)
        self.alpha_max = (min(float(sc.get("alphai", 1.0)) for sc in scalars) ** (-1)) if len(scalars) else 0.0
        self.swork = np.zeros(int(np.prod(self.n)) * (2 if is_impdiff else 1))
        # implicit diffusion: one Helmholtz set-up per velocity component and scalar (main.f90:304-316)
        self.h_vel = []
        if self.is_impdiff:
            for c, cf in enumerate(("fcc", "cfc", "ccf")):
                self.h_vel.append(self.initsolver(cbcvel[6 * c:6 * c + 6], self.bcvel[6 * c:6 * c + 6], cf))
            for sc in self.scalars:
                sc["h"] = self.initsolver(sc["cbc"], sc["bc"], "ccc")

    # initsolver.f90:15-83 (single rank)
    def initsolver(self, cbc, bc, c_or_f):
        n1, n2, n3 = self.n
        lx = eigenvalues(self.ng[0], cbc[0:2], c_or_f[0]) * self.dli[0] ** 2
        ly = eigenvalues(self.ng[1], cbc[2:4], c_or_f[1]) * self.dli[1] ** 2
        lam = np.asfortranarray(lx[:, None] + ly[None, :])
        a, b, c = tridmatrix(cbc[4:6], n3, self.dzci, self.dzfi, c_or_f[2])
        kinds, lens, normfft = fftini(self.ng, cbc[0:4], c_or_f[0:2])
        # initsolver.f90:68-70: the spacings are re-derived from their inverses
        dl = [x ** (-1) for x in self.dli]
        dzc, dzf = self.dzci ** (-1), self.dzfi ** (-1)

        def bcrhs(cb, bv, dlc, dlf, cf):
            out = np.zeros(2)
            dlc_ = np.asarray(dlc, dtype=np.float64)
            dlf_ = np.asarray(dlf, dtype=np.float64)
            lib().orc_bc_rhs(chars(cb), P(np.asarray(bv, dtype=np.float64)), P(dlc_), P(dlf_),
                             chars(cf), P(out))
            return out

        fx = bcrhs(cbc[0:2], bc[0:2], [dl[0], dl[0]], [dl[0], dl[0]], c_or_f[0])
        fy = bcrhs(cbc[2:4], bc[2:4], [dl[1], dl[1]], [dl[1], dl[1]], c_or_f[1])
        if c_or_f[2] == "c":
            fz = bcrhs(cbc[4:6], bc[4:6], [dzc[0], dzc[n3]], [dzf[1], dzf[n3]], "c")
        else:
            fz = bcrhs(cbc[4:6], bc[4:6], [dzc[1], dzc[n3 - 1]], [dzf[1], dzf[n3]], "f")
        rhsbx = np.asfortranarray(np.broadcast_to(fx, (n2, n3, 2)).copy())
        rhsby = np.asfortranarray(np.broadcast_to(fy, (n1, n3, 2)).copy())
        rhsbz = np.asfortranarray(np.broadcast_to(fz, (n1, n2, 2)).copy())
        return dict(lam=lam, a=a, b=b, c=c, kinds=kinds, lens=lens, normfft=normfft,
                    rhsbx=rhsbx, rhsby=rhsby, rhsbz=rhsbz, cbc=cbc, c_or_f=c_or_f)

    # ---- operators (single rank) ----
    def _n(self):
        return IP(self.n)

    def bounduvw(self, u, v, w, keep_norm=-1, halos_norm=-99, halos_tang=-99):
        n, _k = self._n()
        ib, _a = IP(self.is_bound)
        ip, _b = IP(self.is_periodic)
        dl, _c = DP(self.dl)
        lib().orc_bounduvw(chars(self.cbcvel), n, P(self.bcvel), ib, ip, dl, P(self.dzc), P(self.dzf),
                           P(u), P(v), P(w), keep_norm, halos_norm, halos_tang)

    def boundp(self, p, halo_side=-99, cbc=None, bc=None):
        n, _k = self._n()
        ib, _a = IP(self.is_bound)
        ip, _b = IP(self.is_periodic)
        dl, _c = DP(self.dl)
        lib().orc_boundp(chars(cbc or self.cbcpre), n, P(self.bcpre if bc is None else bc), ib, ip, dl,
                         P(self.dzc), P(p), halo_side)

    def fillps(self, u, v, w, pp, dti):
        n, _k = self._n()
        dli, _a = DP(self.dli)
        lib().orc_fillps(n, dli, P(self.dzfi), dti, P(u), P(v), P(w), P(pp))

    def updt_rhs_b(self, pp, s=None, alpha=None):
        s = s or self.p_solver
        n, _k = self._n()
        ib, _a = IP(self.is_bound)
        lib().orc_updt_rhs_b(chars(s["c_or_f"]), chars(s["cbc"]), n, ib, VP(s["rhsbx"]), VP(s["rhsby"]),
                             VP(s["rhsbz"]), P(pp), 0.0 if alpha is None else alpha,
                             0 if alpha is None else 1)

    def solver(self, pp, s=None):
        s = s or self.p_solver
        n, _k = self._n()
        kinds, _a = IP(s["kinds"])
        lens, _b = IP(s["lens"])
        rc = lib().orc_solver(n, kinds, lens, s["normfft"], P(s["lam"]), P(s["a"]), P(s["b"]), P(s["c"]),
                              chars(s["cbc"]), chars(s["c_or_f"]), P(pp))
        if rc:
            raise RuntimeError("orc_solver failed")

    def solve_helmholtz(self, p, alpha, s):
        """solve_helmholtz.f90:28-75 with the arrays of initsolver(cbc, bc, c_or_f): the 3-D branch, or
        the z-only one (solver_gaussel_z) when is_impdiff_1d."""
        n, _k = self._n()
        if self.is_impdiff_1d:
            ib, _c = IP(self.is_bound)
            rc = lib().orc_solve_helmholtz_1d(n, alpha, P(s["a"]), P(s["b"]), P(s["c"]), VP(s["rhsbz"]), ib,
                                              chars(s["cbc"]), chars(s["c_or_f"]), P(p))
            if rc:
                raise RuntimeError("orc_solve_helmholtz_1d failed")
            return
        kinds, _a = IP(s["kinds"])
        lens, _b = IP(s["lens"])
        ib, _c = IP(self.is_bound)
        rc = lib().orc_solve_helmholtz(n, kinds, lens, s["normfft"], alpha, P(s["lam"]), P(s["a"]), P(s["b"]),
                                       P(s["c"]), VP(s["rhsbx"]), VP(s["rhsby"]), VP(s["rhsbz"]), ib,
                                       chars(s["cbc"]), chars(s["c_or_f"]), P(p))
        if rc:
            raise RuntimeError("orc_solve_helmholtz failed")

    def correc(self, pp, u, v, w, dt):
        n, _k = self._n()
        dli, _a = DP(self.dli)
        lib().orc_correc(n, dli, P(self.dzci), dt, P(pp), P(u), P(v), P(w))

    def updatep(self, pp, p, alpha=0.0):
        n, _k = self._n()
        dli, _a = DP(self.dli)
        lib().orc_updatep(n, dli, P(self.dzci), P(self.dzfi), alpha, P(pp), P(p), self.is_impdiff,
                          self.is_impdiff_1d)

    def rk(self, rkpar, dt, p, dudtrko, dvdtrko, dwdtrko, u, v, w, s=None):
        """rk.f90:24-333; s: the first scalar's field when buoyancy is on (rk.f90:69-72)."""
        n, _k = self._n()
        rk_, _a = DP(rkpar)
        dli, _b = DP(self.dli)
        isf, _c = IP(self.is_forced)
        velf, _d = DP(self.velf)
        bfo, _e = DP(self.bforce)
        gacc, _g = DP(self.gacc)
        f = np.zeros(3)
        sb = VP(s) if (s is not None and self.is_buoyancy) else None
        lib().orc_rk(rk_, n, dli, P(self.dzci), P(self.dzfi), P(self.grid_vol_ratio_c),
                     P(self.grid_vol_ratio_f), dt, self.visc, P(p), isf, velf, bfo, gacc, self.beta, sb,
                     P(dudtrko), P(dvdtrko), P(dwdtrko), P(u), P(v), P(w), P(f), P(self.work),
                     self.is_impdiff, self.is_impdiff_1d)
        return f

    def rk_scal(self, rkpar, dt, sc, u, v, w, s, dsdtrko):
        """rk.f90:335-491 for scalar description sc; returns the forcing f."""
        n, _k = self._n()
        rk_, _a = DP(rkpar)
        dli, _b = DP(self.dli)
        f = C.c_double()
        lib().orc_rk_scal(rk_, n, dli, P(self.dzci), P(self.dzfi), P(self.grid_vol_ratio_f), sc["alpha"], dt,
                          P(u), P(v), P(w), sc["is_forced"], sc["scalf"], sc["source"], P(dsdtrko), P(s),
                          C.byref(f), P(self.swork), self.is_impdiff, self.is_impdiff_1d)
        return f.value

    def scal(self, u, v, w, s, alpha):
        """scal.f90:41-98 (explicit diffusion): dsdt as an (n1,n2,n3) array."""
        n, _k = self._n()
        dsdt = np.zeros(tuple(self.n), order="F")
        lib().orc_scal(n, self.dli[0], self.dli[1], P(self.dzci), P(self.dzfi), alpha, P(u), P(v), P(w), P(s),
                       P(dsdt), None, 0, 0)
        return dsdt

    def bulk_forcing_s(self, is_forced, ff, s):
        n, _k = self._n()
        lib().orc_bulk_forcing_s(n, int(is_forced), ff, P(s))

    def chkdiv(self, u, v, w):
        n, _k = self._n()
        l, _a = DP(self.l)
        dli, _b = DP(self.dli)
        tot, mx = C.c_double(), C.c_double()
        lib().orc_chkdiv(n, l, dli, P(self.dzfi), P(u), P(v), P(w), C.byref(tot), C.byref(mx))
        return tot.value, mx.value

    def chkdt(self, u, v, w, alpha=None):
        n, _k = self._n()
        dl, _a = DP(self.dl)
        # alpha_max (param.f90:308-310); with no scalars 1/huge, i.e. max(visc,alpha) = visc
        alpha = self.alpha_max if alpha is None else alpha
        return lib().orc_chkdt(n, dl, P(self.dzci), P(self.dzfi), self.visc, alpha, P(u), P(v), P(w),
                               self.is_impdiff, self.is_impdiff_1d)

    def bulk_mean(self, q, ratio="f"):
        n, _k = self._n()
        g = self.grid_vol_ratio_f if ratio == "f" else self.grid_vol_ratio_c
        return lib().orc_bulk_mean(n, P(g), P(q))

    # ---- main.f90:437-473 (explicit diffusion) ----
    def rk3_step(self, st, dt):
        """One time step on a state dict with u,v,w,p,pp,dudtrko,dvdtrko,dwdtrko (+ s<i>, dsdtrko<i>)."""
        u, v, w, p, pp = st["u"], st["v"], st["w"], st["p"], st["pp"]
        dpdl = np.zeros(3)
        for irk in range(3):
            rkpar = RKCOEFF[irk]
            dtrk = (rkpar[0] + rkpar[1]) * dt
            dtrki = dtrk ** (-1)
            for i, sc in enumerate(self.scalars):  # main.f90:440-450
                s = st["s%d" % i]
                fs = self.rk_scal(rkpar, dt, sc, u, v, w, s, st["dsdtrko%d" % i])
                self.bulk_forcing_s(sc["is_forced"], fs, s)
                if self.is_impdiff:  # main.f90:446-448
                    self.solve_helmholtz(s, -0.5 * sc["alpha"] * dtrk, sc["h"])
                self.boundp(s, cbc=sc["cbc"], bc=sc["bc"])
            f = self.rk(rkpar, dt, p, st["dudtrko"], st["dvdtrko"], st["dwdtrko"], u, v, w,
                        st.get("s0"))
            alpha = 0.0
            if self.is_impdiff:  # main.f90:453-462
                alpha = -.5 * self.visc * dtrk
                for q, h in zip((u, v, w), self.h_vel):
                    self.solve_helmholtz(q, alpha, h)
            dpdl += f
            self.bounduvw(u, v, w, 0, -1)
            self.fillps(u, v, w, pp, dtrki)
            self.updt_rhs_b(pp)
            self.solver(pp)
            self.boundp(pp)
            self.correc(pp, u, v, w, dtrk)
            self.bounduvw(u, v, w, 1, 1, 0)
            self.updatep(pp, p, alpha)
            self.boundp(p, 1)
        return dpdl

    def new_state(self):
        n = self.n
        st = {k: field(n) for k in ("u", "v", "w", "p", "pp")}
        for k in ("dudtrko", "dvdtrko", "dwdtrko"):
            st[k] = np.zeros(tuple(n), order="F")
        for i in range(len(self.scalars)):
            st["s%d" % i] = field(n)
            st["dsdtrko%d" % i] = np.zeros(tuple(n), order="F")
        return st

    def init_bounds(self, st):
        """main.f90:389-390 (and the scalars' boundp of main.f90:391-395)."""
        self.bounduvw(st["u"], st["v"], st["w"])
        self.boundp(st["p"], 1)
        for i, sc in enumerate(self.scalars):
            self.boundp(st["s%d" % i], cbc=sc["cbc"], bc=sc["bc"])
