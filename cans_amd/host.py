"""Host-side mirror of the reference's set-up and time loop over the HIP C-ABI.

The reference's Fortran host (src/main.f90, src/initsolver.f90, src/param.f90)
computes the grid and solver coefficients on the CPU and drives the per-routine
device calls.  This module is the same host in Python: ``Case`` holds the
namelist parameters of src/param.f90, ``Simulation`` performs the
initialisation of src/main.f90:241-395 and the RK3 substeps of
src/main.f90:437-473 through ``libcans_hip.so``.  Device memory comes from
PyTorch tensors laid out as the context's halo'd fields.
"""
import ctypes as C
from dataclasses import dataclass, field

import numpy as np

from . import _lib as LB

RKCOEFF = ((32. / 60., 0.), (25. / 60., -17. / 60.), (45. / 60., -25. / 60.))  # src/param.f90:20-22


def _torch():
    import torch
    return torch


@dataclass
class Case:
    """Parameters of the &dns namelist (src/param.f90:27-105) that the step uses."""
    ng: tuple
    l: tuple
    gtype: int = 1
    gr: float = 0.0
    # &numerics is_gridpoint_natural_channel (param.f90:99): gridpoint_natural faces, gtype/gr unused
    is_gridpoint_natural_channel: bool = False
    cbcvel: str = "PP" * 9          # cbcvel(0:1,3,3) flattened, component slowest
    cbcpre: str = "PP" * 3          # cbcpre(0:1,3)
    bcvel: tuple = (0.0,) * 18
    bcpre: tuple = (0.0,) * 6
    visci: float = 1.0
    bforce: tuple = (0.0, 0.0, 0.0)
    is_forced: tuple = (False, False, False)
    velf: tuple = (0.0, 0.0, 0.0)
    # &scalar namelist (src/param.f90:140-152), one dict per scalar with keys alphai, cbc
    # (6 chars, cbcscal(0:1,1:3)), bc (6), source, is_forced, scalf; gacc/beta of &dns
    scalars: tuple = ()
    gacc: tuple = (0.0, 0.0, 0.0)
    beta: float = 0.0
    is_boussinesq_buoyancy: bool = False
    # &numerics (src/param.f90:96,331): implicit diffusion, in 3-D or in z only
    is_impdiff: bool = False
    is_impdiff_1d: bool = False
    # is_poisson_dtdma: z slabs of the distributed tridiagonal solve (0: the plain sweep)
    dtdma_slabs: int = 0

    @property
    def n(self):
        return tuple(int(x) for x in self.ng)


class Grid:
    """dl, dli, z-grid and volume ratios (src/param.f90:196-198, src/main.f90:241-285)."""

    def __init__(self, case: Case):
        L = LB.lib()
        ng = case.n
        self.n = ng
        self.l = tuple(float(x) for x in case.l)
        # (1.*ng) is a default-real (single precision) expression in param.f90:196
        self.dl = tuple(self.l[d] / float(np.float32(ng[d])) for d in range(3))
        self.dli = tuple(self.dl[d] ** (-1) for d in range(3))
        self.visc = float(case.visci) ** (-1)
        n3 = ng[2]
        self.dzc, self.dzf, self.zc, self.zf = (np.zeros(n3 + 2) for _ in range(4))
        periodic_z = case.cbcpre[4:6] == "PP"
        if case.is_gridpoint_natural_channel:
            LB.check(L.cans_initgrid_natural(n3, self.l[2], int(periodic_z), _dp(self.dzc), _dp(self.dzf),
                                             _dp(self.zc), _dp(self.zf)), "cans_initgrid_natural")
        else:
            LB.check(L.cans_initgrid(case.gtype, n3, case.gr, self.l[2], int(periodic_z), _dp(self.dzc),
                                     _dp(self.dzf), _dp(self.zc), _dp(self.zf)), "cans_initgrid")
        self.dzci = self.dzc ** (-1)
        self.dzfi = self.dzf ** (-1)
        vol = self.l[0] * self.l[1] * self.l[2]
        self.grid_vol_ratio_c = self.dl[0] * self.dl[1] * self.dzc / vol
        self.grid_vol_ratio_f = self.dl[0] * self.dl[1] * self.dzf / vol
        self.is_periodic = tuple(case.cbcpre[2 * d] == "P" and case.cbcpre[2 * d + 1] == "P" for d in range(3))
        # one GPU holds the whole domain as an x-pencil (src/initmpi.f90:226-267):
        # x is always a bounded direction, y/z are bounded unless periodic;
        # periodic y/z have this rank as neighbour (nb >= 0), others MPI_PROC_NULL (-1)
        self.is_bound = (1, 1) + tuple(0 if self.is_periodic[d] else 1 for d in (1, 1, 2, 2))
        self.nb = (-1, -1) + tuple(0 if self.is_periodic[d] else -1 for d in (1, 1, 2, 2))


def _dp(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


def eigenvalues(n, bc, c_or_f):
    lam = np.zeros(n)
    LB.check(LB.lib().cans_eigenvalues(n, bc.encode(), c_or_f.encode(), _dp(lam)), "cans_eigenvalues")
    return lam


def bc_rhs(cbc, bc, dlc, dlf, c_or_f):
    """initsolver.f90:189-232: the two plane values of the boundary rhs."""
    out = np.zeros(2)
    LB.check(LB.lib().cans_bc_rhs(cbc.encode(), _dp(np.ascontiguousarray(bc, dtype=np.float64)),
                                  _dp(np.ascontiguousarray(dlc, dtype=np.float64)),
                                  _dp(np.ascontiguousarray(dlf, dtype=np.float64)), c_or_f.encode(), _dp(out)),
             "cans_bc_rhs")
    return out


def tridmatrix(bc, n, dzci, dzfi, c_or_f):
    a, b, c = np.zeros(n), np.zeros(n), np.zeros(n)
    LB.check(LB.lib().cans_tridmatrix(bc.encode(), n, _dp(np.ascontiguousarray(dzci)),
                                      _dp(np.ascontiguousarray(dzfi)), c_or_f.encode(), _dp(a), _dp(b), _dp(c)),
             "cans_tridmatrix")
    return a, b, c


class Context:
    """cans_ctx: one grid, one HIP stream, one precision mode.

    ``precision`` is ``CANS_FP64`` (every array double) or ``CANS_MIXED``
    (velocity double, pressure-like arrays float; every kernel computes in
    double)."""

    def __init__(self, n, ld=None, precision=LB.CANS_FP64, stream=None, device="cuda"):
        """stream: None (torch's current stream), or a torch.cuda.Stream the context owns
        for its lifetime (required to capture its work in a HIP graph)."""
        torch = _torch()
        self.torch = torch
        self.device = torch.device(device)
        torch.cuda.init()
        self.n = tuple(int(x) for x in n)
        self.precision = int(precision)
        # storage type per field kind (include/cans_hip.h: velocity / pressure)
        mixed = self.precision == LB.CANS_MIXED
        self.dtypes = {LB.CANS_FIELD_VELOCITY: torch.float64,
                       LB.CANS_FIELD_PRESSURE: torch.float32 if mixed else torch.float64}
        self.p_dtype = self.dtypes[LB.CANS_FIELD_PRESSURE]
        self.p_es = 4 if mixed else 8
        if ld is None:  # interior rows start 256-byte aligned
            ld = ((self.n[0] + 2 + 31) // 32) * 32
        self.torch_stream = stream if stream is not None else torch.cuda.current_stream(self.device)
        stream = self.torch_stream.cuda_stream
        self.stream = stream
        h = C.c_void_p()
        LB.check(LB.lib().cans_ctx_create(LB.ivec(self.n), int(ld), int(precision), C.c_void_p(stream),
                                          C.byref(h)), "cans_ctx_create")
        self.h = h
        sy, sz, el = C.c_long(), C.c_long(), C.c_long()
        LB.check(LB.lib().cans_ctx_layout(h, C.byref(sy), C.byref(sz), C.byref(el)), "cans_ctx_layout")
        self.sy, self.sz, self.elems = sy.value, sz.value, el.value

    def close(self):
        if self.h:
            LB.lib().cans_ctx_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def sync(self):
        LB.check(LB.lib().cans_ctx_sync(self.h), "cans_ctx_sync")

    # halo'd field backed by a torch tensor; .ptr is the address of element (0,0,0)
    def field(self, fill=0.0, kind=LB.CANS_FIELD_VELOCITY):
        return DeviceField(self, fill, kind)

    def darray(self, a, dtype=None):
        torch = self.torch
        return torch.as_tensor(np.ascontiguousarray(a), dtype=dtype or torch.float64, device=self.device)


class DeviceField:
    """Halo'd field of one kind (velocity or pressure) in the context's storage type for it."""

    def __init__(self, ctx: Context, fill=0.0, kind=LB.CANS_FIELD_VELOCITY):
        torch = ctx.torch
        self.ctx = ctx
        self.dtype = ctx.dtypes[kind]
        es = 8 if self.dtype == torch.float64 else 4
        self.buf = torch.full((ctx.elems + 2 * 256 // es,), float(fill), dtype=self.dtype, device=ctx.device)
        base = self.buf.data_ptr()
        # element 1 of every row on a 256-byte boundary
        shift = ((-(base + es)) % 256) // es
        self.off = shift
        self.ptr = base + es * shift
        n1, n2, n3 = ctx.n
        self.view = self.buf[shift:shift + ctx.elems].view(n3 + 2, n2 + 2, ctx.sy)[:, :, :n1 + 2]
        # tell the library the element type behind this address (see cans_field_register)
        LB.check(LB.lib().cans_field_register(ctx.h, C.c_void_p(self.ptr), kind), "cans_field_register")

    def numpy(self):
        """Fortran-ordered (n1+2, n2+2, n3+2) float64 copy on the host (FP32 storage widens exactly)."""
        return np.asfortranarray(self.view.cpu().numpy().astype(np.float64).transpose(2, 1, 0))

    def set(self, a):
        """Store a float64 host array (rounded to nearest for FP32 storage)."""
        a = np.asarray(a, dtype=np.float64)
        n1, n2, n3 = self.ctx.n
        assert a.shape == (n1 + 2, n2 + 2, n3 + 2)
        # pad rows to the leading dimension on the host, then one contiguous transfer
        sy = self.ctx.sy
        host = np.zeros((n3 + 2, n2 + 2, sy), dtype=np.float32 if self.dtype != self.ctx.torch.float64 else np.float64)
        host[:, :, :n1 + 2] = a.transpose(2, 1, 0)
        t = self.ctx.torch.from_numpy(host.reshape(-1))
        self.buf[self.off:self.off + host.size].copy_(t)

    def interior(self):
        return self.numpy()[1:-1, 1:-1, 1:-1]


class Plan:
    """Transforms + solver arrays of one variable (src/initsolver.f90:15-83)."""

    def __init__(self, ctx: Context, grid: Grid, cbc: str, bc, c_or_f="ccc", transforms=True):
        self.grid = grid
        self.ctx = ctx
        self.transforms = transforms  # False: z-only systems (is_impdiff_1d), no x/y plans
        ng = grid.n
        lx = eigenvalues(ng[0], cbc[0:2], c_or_f[0]) * grid.dli[0] ** 2
        ly = eigenvalues(ng[1], cbc[2:4], c_or_f[1]) * grid.dli[1] ** 2
        self.lam_h = np.asfortranarray(lx[:, None] + ly[None, :])  # lambdaxy(i,j), device ordering
        self.a_h, self.b_h, self.c_h = tridmatrix(cbc[4:6], ng[2], grid.dzci, grid.dzfi, c_or_f[2])
        h = C.c_void_p()
        nf = C.c_double(1.0)
        if transforms:
            LB.check(LB.lib().cans_fftini(ctx.h, LB.ivec(ng), cbc[0:4].encode(), c_or_f[0:2].encode(), C.byref(h),
                                          C.byref(nf)), "cans_fftini")
        self.h = h
        self.normfft = nf.value
        self.cbc = cbc
        self.c_or_f = c_or_f
        self.lam = ctx.darray(self.lam_h.ravel(order="F"))
        self.a = ctx.darray(self.a_h)
        self.b = ctx.darray(self.b_h)
        self.c = ctx.darray(self.c_h)
        # boundary rhs planes (initsolver.f90:66-77): spacings re-derived from their inverses
        bc = np.asarray(bc, dtype=np.float64)
        n1, n2, n3 = ng
        dl = [x ** (-1) for x in grid.dli]
        dzc, dzf = grid.dzci ** (-1), grid.dzfi ** (-1)
        fx = bc_rhs(cbc[0:2], bc[0:2], [dl[0], dl[0]], [dl[0], dl[0]], c_or_f[0])
        fy = bc_rhs(cbc[2:4], bc[2:4], [dl[1], dl[1]], [dl[1], dl[1]], c_or_f[1])
        if c_or_f[2] == "c":
            fz = bc_rhs(cbc[4:6], bc[4:6], [dzc[0], dzc[n3]], [dzf[1], dzf[n3]], "c")
        else:
            fz = bc_rhs(cbc[4:6], bc[4:6], [dzc[1], dzc[n3 - 1]], [dzf[1], dzf[n3]], "f")
        # rhsbx(n2,n3,0:1), rhsby(n1,n3,0:1), rhsbz(n1,n2,0:1): constant planes, Fortran order
        self.rhsb = [ctx.darray(np.repeat(f, m)) for f, m in ((fx, n2 * n3), (fy, n1 * n3), (fz, n1 * n2))]

    def close(self):
        if self.h:
            LB.lib().cans_fftend(self.h)
            self.h = None

    def rhsb_ptrs(self):
        """Boundary rhs planes; the z-only form has no x/y ones (main.f90:317-329)."""
        ptrs = [t.data_ptr() for t in self.rhsb]
        return ptrs if self.transforms else [None, None, ptrs[2]]

    def solve_helmholtz(self, f: DeviceField, alpha):
        """solve_helmholtz.f90:28-75 on f (the rhs in, the solution out)."""
        n = LB.ivec(self.ctx.n)
        LB.check(LB.lib().cans_solve_helmholtz(self.ctx.h, n, n, self.h, self.normfft, alpha,
                                               *[C.c_void_p(t.data_ptr()) for t in (self.lam, self.a, self.b, self.c)],
                                               *[C.c_void_p(x) for x in self.rhsb_ptrs()],
                                               LB.ivec(self.grid.is_bound), self.cbc.encode(), self.c_or_f.encode(),
                                               C.c_void_p(f.ptr)), "cans_solve_helmholtz")

    def fill(self, h):
        """Fill a cans_helmholtz struct with this system."""
        h.plan = self.h.value if self.h else None
        h.normfft = self.normfft
        h.lambdaxy_d, h.a_d, h.b_d, h.c_d = (t.data_ptr() for t in (self.lam, self.a, self.b, self.c))
        h.rhsbx_d, h.rhsby_d, h.rhsbz_d = self.rhsb_ptrs()
        h.cbc = self.cbc.encode()
        h.c_or_f = self.c_or_f.encode()

    def solve(self, f: DeviceField):
        n = LB.ivec(self.ctx.n)
        LB.check(LB.lib().cans_solver(self.ctx.h, n, n, self.h, self.normfft, C.c_void_p(self.lam.data_ptr()),
                                      C.c_void_p(self.a.data_ptr()), C.c_void_p(self.b.data_ptr()),
                                      C.c_void_p(self.c.data_ptr()), self.cbc.encode(), self.c_or_f.encode(),
                                      C.c_void_p(f.ptr)), "cans_solver")


class PoissonBatch:
    """Independent Poisson problems of different sizes solved together
    (cans_poisson_batch_*): per box, the arrays of initsolver (src/initsolver.f90:15-83,
    c_or_f = 'ccc') and one src/solver.f90:17 call, all boxes in five launches.

    boxes: dicts with n=(n1, n2, n3), l=(lx, ly, lz), cbc (cbcpre, 6 chars, z
    non-periodic) and optionally gtype (default 1) and gr (default 0, uniform z).
    Each box's field is halo'd, dense (leading dimension n1+2), in the context's
    pressure storage type."""

    def __init__(self, ctx: Context, boxes):
        torch = ctx.torch
        self.ctx = ctx
        es = ctx.p_es
        self.n, self.views, self.bufs, self.dzfi = [], [], [], []
        arr = (LB.PoissonBox * len(boxes))()
        keep = []
        for i, bx in enumerate(boxes):
            n1, n2, n3 = (int(v) for v in bx["n"])
            lx, ly, lz = (float(v) for v in bx["l"])
            cbc = bx["cbc"]
            dli = [(x / float(np.float32(m))) ** (-1) for x, m in ((lx, n1), (ly, n2))]
            dzc, dzf, zc, zf = (np.zeros(n3 + 2) for _ in range(4))
            LB.check(LB.lib().cans_initgrid(int(bx.get("gtype", 1)), n3, float(bx.get("gr", 0.0)), lz, 0, _dp(dzc),
                                            _dp(dzf), _dp(zc), _dp(zf)), "cans_initgrid")
            dzci, dzfi = dzc ** (-1), dzf ** (-1)
            lam = (eigenvalues(n1, cbc[0:2], "c")[:, None] * dli[0] ** 2 +
                   eigenvalues(n2, cbc[2:4], "c")[None, :] * dli[1] ** 2).ravel(order="F")
            a, b, c = tridmatrix(cbc[4:6], n3, dzci, dzfi, "c")
            sy, sz = n1 + 2, (n1 + 2) * (n2 + 2)
            elems = sz * (n3 + 2)
            buf = torch.zeros(elems + 2 * 256 // es, dtype=ctx.p_dtype, device=ctx.device)
            shift = ((-(buf.data_ptr() + es)) % 256) // es  # element (1,0,0) on a 256-byte boundary
            self.bufs.append(buf)
            self.views.append(buf[shift:shift + elems].view(n3 + 2, n2 + 2, sy))
            self.n.append((n1, n2, n3))
            self.dzfi.append(dzfi)
            arrays = [np.ascontiguousarray(q, dtype=np.float64) for q in (lam, a, b, c)]
            keep.append(arrays)
            bs = arr[i]
            bs.n[:] = [n1, n2, n3]
            bs.sy, bs.sz = sy, sz
            bs.cbc = cbc.encode()
            bs.lambdaxy_h, bs.a_h, bs.b_h, bs.c_h = (q.ctypes.data for q in arrays)
            bs.p_d = buf.data_ptr() + es * shift
        nf = np.zeros(len(boxes))
        h = C.c_void_p()
        LB.check(LB.lib().cans_poisson_batch_create(ctx.h, len(boxes), arr, _dp(nf), C.byref(h)),
                 "cans_poisson_batch_create")
        self.h = h
        self.normfft = nf

    def close(self):
        if self.h:
            LB.lib().cans_poisson_batch_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set(self, i, a):
        """Store box i from a Fortran-shaped (n1+2, n2+2, n3+2) float64 host array."""
        a = np.asarray(a, dtype=np.float64)
        assert a.shape == tuple(m + 2 for m in self.n[i])
        host = np.ascontiguousarray(a.transpose(2, 1, 0)).astype(np.float32 if self.ctx.p_es == 4 else np.float64)
        self.views[i].copy_(self.ctx.torch.from_numpy(host))

    def get(self, i):
        """Box i as a Fortran-ordered (n1+2, n2+2, n3+2) float64 host array."""
        return np.asfortranarray(self.views[i].cpu().numpy().astype(np.float64).transpose(2, 1, 0))

    def solve(self):
        LB.check(LB.lib().cans_poisson_batch_solve(self.h), "cans_poisson_batch_solve")

    def threads(self):
        """Lanes per x/y work-group the batch's tuner chose."""
        t = C.c_int()
        LB.check(LB.lib().cans_poisson_batch_tuning(self.h, C.byref(t)), "cans_poisson_batch_tuning")
        return t.value


class Simulation:
    """One case on one GPU: set-up of src/main.f90:241-395 and its RK3 loop."""

    def __init__(self, case: Case, ctx: Context = None, precision=LB.CANS_FP64, stream=None):
        self.case = case
        self.grid = Grid(case)
        self.ctx = ctx or Context(case.n, precision=precision, stream=stream)
        ctx = self.ctx
        g = self.grid
        self.imp = 2 if case.is_impdiff_1d else 1 if case.is_impdiff else 0
        LB.check(LB.lib().cans_ctx_set_numerics(ctx.h, int(self.imp > 0), int(self.imp == 2)),
                 "cans_ctx_set_numerics")
        LB.check(LB.lib().cans_ctx_set_dtdma(ctx.h, int(case.dtdma_slabs)), "cans_ctx_set_dtdma")
        self.plan = Plan(ctx, g, case.cbcpre, case.bcpre, "ccc")
        # implicit diffusion: one Helmholtz system per velocity component (main.f90:304-309)
        self.hplans = []
        if self.imp:
            for c, cf in enumerate(("fcc", "cfc", "ccf")):
                self.hplans.append(Plan(ctx, g, case.cbcvel[6 * c:6 * c + 6], case.bcvel[6 * c:6 * c + 6], cf,
                                        transforms=self.imp == 1))
        self.u, self.v, self.w = (ctx.field() for _ in range(3))
        self.p, self.pp = (ctx.field(kind=LB.CANS_FIELD_PRESSURE) for _ in range(2))
        torch = ctx.torch
        n1, n2, n3 = case.n
        self.rhs_o = [torch.zeros(n1 * n2 * n3, dtype=ctx.p_dtype, device=ctx.device) for _ in range(3)]
        self.dzci_d = ctx.darray(g.dzci)
        self.dzfi_d = ctx.darray(g.dzfi)
        self.gvr_c_d = ctx.darray(g.grid_vol_ratio_c)
        self.gvr_f_d = ctx.darray(g.grid_vol_ratio_f)
        self._dzc = np.ascontiguousarray(g.dzc)
        self._dzf = np.ascontiguousarray(g.dzf)
        prm = LB.StepParams()
        prm.cbcvel = case.cbcvel.encode()
        prm.bcvel = LB.dvec(case.bcvel)
        prm.cbcpre = case.cbcpre.encode()
        prm.bcpre = LB.dvec(case.bcpre)
        prm.nb = LB.ivec(g.nb)
        prm.is_bound = LB.ivec(g.is_bound)
        prm.dl = LB.dvec(g.dl)
        prm.dli = LB.dvec(g.dli)
        prm.visc = g.visc
        prm.is_forced = LB.ivec([int(bool(x)) for x in case.is_forced])
        prm.velf = LB.dvec(case.velf)
        prm.bforce = LB.dvec(case.bforce)
        prm.dzc_h = self._dzc.ctypes.data
        prm.dzf_h = self._dzf.ctypes.data
        prm.dzci_d = self.dzci_d.data_ptr()
        prm.dzfi_d = self.dzfi_d.data_ptr()
        prm.grid_vol_ratio_c_d = self.gvr_c_d.data_ptr()
        prm.grid_vol_ratio_f_d = self.gvr_f_d.data_ptr()
        prm.plan = self.plan.h.value
        prm.normfft = self.plan.normfft
        prm.lambdaxy_d = self.plan.lam.data_ptr()
        prm.a_d = self.plan.a.data_ptr()
        prm.b_d = self.plan.b.data_ptr()
        prm.c_d = self.plan.c.data_ptr()
        prm.rhsbx_d = prm.rhsby_d = prm.rhsbz_d = None  # homogeneous pressure BCs in the solved cases
        if any(abs(x) > 0 for x in case.bcpre):
            raise NotImplementedError("non-homogeneous pressure boundary values")
        self.prm = prm
        # scalars: FP64 fields in every precision mode (scal.f90:17-41)
        self.scal_spec = [dict(alpha=float(sc.get("alphai", 1.0)) ** (-1), cbc=sc["cbc"],
                               bc=tuple(float(x) for x in sc.get("bc", (0.0,) * 6)),
                               source=float(sc.get("source", 0.0)), is_forced=int(bool(sc.get("is_forced", False))),
                               scalf=float(sc.get("scalf", 0.0))) for sc in case.scalars]
        if len(self.scal_spec) > 32:
            raise ValueError("at most CANS_MAX_SCALARS (32) scalars")
        self.s = [ctx.field() for _ in self.scal_spec]
        self.dsdtrko = [torch.zeros(n1 * n2 * n3, dtype=torch.float64, device=ctx.device) for _ in self.scal_spec]
        # param.f90:308-310: the largest scalar diffusivity for chkdt (1/huge without scalars)
        self.alpha_max = max((sp["alpha"] for sp in self.scal_spec), default=0.0)
        self.is_buoyancy = bool(case.is_boussinesq_buoyancy) and len(self.s) > 0
        self._scal_c = (LB.Scalar * max(1, len(self.s)))()
        self.splans = [Plan(ctx, g, sp["cbc"], sp["bc"], "ccc", transforms=self.imp == 1) if self.imp else None
                       for sp in self.scal_spec]
        for i, sp in enumerate(self.scal_spec):
            e = self._scal_c[i]
            e.val, e.dsdtrko = self.s[i].ptr, self.dsdtrko[i].data_ptr()
            e.alpha, e.cbc, e.bc = sp["alpha"], sp["cbc"].encode(), LB.dvec(sp["bc"])
            e.source, e.is_forced, e.scalf = sp["source"], sp["is_forced"], sp["scalf"]
            if self.imp:
                self.splans[i].fill(e.h)
        for k, hp in enumerate(self.hplans):
            hp.fill(prm.hvel[k])
        prm.nscal = len(self.s)
        prm.scal = C.cast(self._scal_c, C.c_void_p).value if self.s else None
        prm.is_buoyancy = int(self.is_buoyancy)
        prm.gacc = LB.dvec(case.gacc)
        prm.beta = float(case.beta)
        st = LB.State()
        st.u, st.v, st.w, st.p, st.pp = (f.ptr for f in (self.u, self.v, self.w, self.p, self.pp))
        st.dudtrko, st.dvdtrko, st.dwdtrko = (t.data_ptr() for t in self.rhs_o)
        self.st = st

    def set_refinement(self, sweeps, substep_mask=4):
        """cans_ctx_set_refinement: FP64 refinement sweeps of the fused mixed substep's pressure
        solve on the substeps of substep_mask (bit irk-1; 4 = the last substep of each step)."""
        LB.check(LB.lib().cans_ctx_set_refinement(self.ctx.h, int(sweeps), int(substep_mask)),
                 "cans_ctx_set_refinement")

    def set_mixed_increments(self, on=True):
        """cans_ctx_set_mixed_increments: keep the mixed step's intermediate velocity as FP32
        increments over the step's velocity instead of an FP64 field."""
        LB.check(LB.lib().cans_ctx_set_mixed_increments(self.ctx.h, int(bool(on))), "cans_ctx_set_mixed_increments")

    # ---- state ----
    def set_state(self, u=None, v=None, w=None, p=None, s=()):
        for f, a in ((self.u, u), (self.v, v), (self.w, w), (self.p, p)) + tuple(zip(self.s, s)):
            if a is not None:
                f.set(a)

    def init_bounds(self, restart=False):
        """main.f90:389-395."""
        self.bounduvw(self.u, self.v, self.w, keep_norm=1 if restart else 0)
        self.boundp(self.p, LB.HALO_UPPER)
        for i in range(len(self.s)):
            self.boundp_s(i)

    # ---- operators (reference argument lists; single GPU) ----
    def bounduvw(self, u, v, w, keep_norm=-1, halos_norm=LB.HALO_ABSENT, halos_tang=LB.HALO_ABSENT):
        c, g = self.case, self.grid
        LB.check(LB.lib().cans_bounduvw(self.ctx.h, c.cbcvel.encode(), LB.ivec(c.n), LB.dvec(c.bcvel),
                                        LB.ivec(g.nb), LB.ivec(g.is_bound), LB.dvec(g.dl), _dp(self._dzc),
                                        _dp(self._dzf), C.c_void_p(u.ptr), C.c_void_p(v.ptr), C.c_void_p(w.ptr),
                                        keep_norm, halos_norm, halos_tang), "cans_bounduvw")

    def boundp(self, p, halo_side=LB.HALO_ABSENT):
        c, g = self.case, self.grid
        LB.check(LB.lib().cans_boundp(self.ctx.h, c.cbcpre.encode(), LB.ivec(c.n), LB.dvec(c.bcpre),
                                      LB.ivec(g.nb), LB.ivec(g.is_bound), LB.dvec(g.dl), _dp(self._dzc),
                                      C.c_void_p(p.ptr), halo_side), "cans_boundp")

    def rk(self, irk, dt, with_f=True):
        c, g = self.case, self.grid
        f = (C.c_double * 3)()
        sb = C.c_void_p(self.s[0].ptr) if self.is_buoyancy else None
        LB.check(LB.lib().cans_rk(self.ctx.h, LB.dvec(RKCOEFF[irk - 1]), LB.ivec(c.n), LB.dvec(g.dli),
                                  C.c_void_p(self.dzci_d.data_ptr()), C.c_void_p(self.dzfi_d.data_ptr()),
                                  C.c_void_p(self.gvr_c_d.data_ptr()), C.c_void_p(self.gvr_f_d.data_ptr()),
                                  dt, g.visc, C.c_void_p(self.p.ptr),
                                  LB.ivec([int(bool(x)) for x in c.is_forced]), LB.dvec(c.velf),
                                  LB.dvec(c.bforce), LB.dvec(c.gacc), float(c.beta), sb,
                                  *[C.c_void_p(t.data_ptr()) for t in self.rhs_o],
                                  C.c_void_p(self.u.ptr), C.c_void_p(self.v.ptr), C.c_void_p(self.w.ptr),
                                  f if with_f else None), "cans_rk")
        return np.array(f[:]) if with_f else None

    def rk_scal(self, i, irk, dt):
        """rk.f90:335 for scalar i; returns its bulk forcing f."""
        c, g, sp = self.case, self.grid, self.scal_spec[i]
        f = C.c_double()
        fluxo = (C.c_double * 6)()
        LB.check(LB.lib().cans_rk_scal(self.ctx.h, LB.dvec(RKCOEFF[irk - 1]), LB.ivec(c.n), LB.dvec(g.dli),
                                       LB.dvec(g.l), C.c_void_p(self.dzci_d.data_ptr()),
                                       C.c_void_p(self.dzfi_d.data_ptr()), C.c_void_p(self.gvr_f_d.data_ptr()),
                                       sp["alpha"], dt, LB.ivec(g.is_bound), C.c_void_p(self.u.ptr),
                                       C.c_void_p(self.v.ptr), C.c_void_p(self.w.ptr), sp["is_forced"], sp["scalf"],
                                       sp["source"], fluxo, C.c_void_p(self.dsdtrko[i].data_ptr()),
                                       C.c_void_p(self.s[i].ptr), C.byref(f)), "cans_rk_scal")
        return f.value

    def bulk_forcing_s(self, i, ff):
        LB.check(LB.lib().cans_bulk_forcing_s(self.ctx.h, LB.ivec(self.case.n), self.scal_spec[i]["is_forced"], ff,
                                              C.c_void_p(self.s[i].ptr)), "cans_bulk_forcing_s")

    def boundp_s(self, i):
        """boundp of scalar i with its own cbc/bc (main.f90:449)."""
        c, g, sp = self.case, self.grid, self.scal_spec[i]
        LB.check(LB.lib().cans_boundp(self.ctx.h, sp["cbc"].encode(), LB.ivec(c.n), LB.dvec(sp["bc"]),
                                      LB.ivec(g.nb), LB.ivec(g.is_bound), LB.dvec(g.dl), _dp(self._dzc),
                                      C.c_void_p(self.s[i].ptr), LB.HALO_ABSENT), "cans_boundp")

    def fillps(self, dti):
        g = self.grid
        LB.check(LB.lib().cans_fillps(self.ctx.h, LB.ivec(g.n), LB.dvec(g.dli), C.c_void_p(self.dzfi_d.data_ptr()),
                                      dti, C.c_void_p(self.u.ptr), C.c_void_p(self.v.ptr), C.c_void_p(self.w.ptr),
                                      C.c_void_p(self.pp.ptr)), "cans_fillps")

    def correc(self, dt):
        g = self.grid
        LB.check(LB.lib().cans_correc(self.ctx.h, LB.ivec(g.n), LB.dvec(g.dli), C.c_void_p(self.dzci_d.data_ptr()),
                                      dt, C.c_void_p(self.pp.ptr), C.c_void_p(self.u.ptr), C.c_void_p(self.v.ptr),
                                      C.c_void_p(self.w.ptr)), "cans_correc")

    def updatep(self, alpha=0.0):
        g = self.grid
        LB.check(LB.lib().cans_updatep(self.ctx.h, LB.ivec(g.n), LB.dvec(g.dli),
                                       C.c_void_p(self.dzci_d.data_ptr()), C.c_void_p(self.dzfi_d.data_ptr()), alpha,
                                       C.c_void_p(self.pp.ptr), C.c_void_p(self.p.ptr), int(self.imp > 0)),
                 "cans_updatep")

    def solver(self, f=None):
        self.plan.solve(f or self.pp)

    def solver_refined(self, f, sweeps):
        """cans_solver_refined on an FP64 DeviceField f (velocity kind): rhs in, solution out; the inner
        solves in the context's pressure storage type, the residual and the sum in FP64."""
        g, pl = self.grid, self.plan
        n = LB.ivec(g.n)
        LB.check(LB.lib().cans_solver_refined(self.ctx.h, n, n, pl.h, pl.normfft,
                                              *[C.c_void_p(t.data_ptr()) for t in (pl.lam, pl.a, pl.b, pl.c)],
                                              pl.cbc.encode(), pl.c_or_f.encode(), LB.dvec(g.dli[:2]),
                                              C.c_void_p(self.dzci_d.data_ptr()), C.c_void_p(self.dzfi_d.data_ptr()),
                                              int(sweeps), C.c_void_p(f.ptr)), "cans_solver_refined")

    def chkdt(self):
        g = self.grid
        out = C.c_double()
        LB.check(LB.lib().cans_chkdt(self.ctx.h, LB.ivec(g.n), LB.dvec(g.dl), C.c_void_p(self.dzci_d.data_ptr()),
                                     C.c_void_p(self.dzfi_d.data_ptr()), g.visc, self.alpha_max,
                                     C.c_void_p(self.u.ptr),
                                     C.c_void_p(self.v.ptr), C.c_void_p(self.w.ptr), int(self.imp > 0),
                                     int(self.imp == 2), C.byref(out)),
                 "cans_chkdt")
        return out.value

    def chkdiv(self):
        g = self.grid
        tot, mx = C.c_double(), C.c_double()
        LB.check(LB.lib().cans_chkdiv(self.ctx.h, LB.ivec((1, 1, 1)), LB.ivec(g.n), LB.dvec(g.l), LB.dvec(g.dli),
                                      C.c_void_p(self.dzfi_d.data_ptr()), C.c_void_p(self.u.ptr),
                                      C.c_void_p(self.v.ptr), C.c_void_p(self.w.ptr), C.byref(tot), C.byref(mx)),
                 "cans_chkdiv")
        return tot.value, mx.value

    # ---- main.f90:437-473 ----
    def rk3_substep(self, irk, dt):
        LB.check(LB.lib().cans_rk3_substep(self.ctx.h, C.byref(self.prm), irk, dt, C.byref(self.st)),
                 "cans_rk3_substep")

    def rk3_step(self, dt):
        """cans_rk3_step: one whole step (the substeps' projections folded where that applies)."""
        LB.check(LB.lib().cans_rk3_step(self.ctx.h, C.byref(self.prm), dt, C.byref(self.st)), "cans_rk3_step")

    def solve_helmholtz_batch(self, fields, plans, alphas):
        """cans_solve_helmholtz_batch: solve_helmholtz (solve_helmholtz.f90:28-75) of each FP64
        DeviceField with its Plan and alpha, as one batch."""
        n = len(fields)
        hs = (LB.Helmholtz * n)()
        for q, pl in enumerate(plans):
            pl.fill(hs[q])
        al = (C.c_double * n)(*alphas)
        fp = (C.c_void_p * n)(*[f.ptr for f in fields])
        LB.check(LB.lib().cans_solve_helmholtz_batch(self.ctx.h, n, hs, al, LB.ivec(self.grid.is_bound), fp),
                 "cans_solve_helmholtz_batch")

    def rk3_step_substeps(self, dt):
        """The same step as three cans_rk3_substep calls."""
        for irk in (1, 2, 3):
            self.rk3_substep(irk, dt)

    def rk3_substep_adjoint(self, irk, dt, q, ub, fb, ubo, fbo, pp, scal=None):
        """cans_rk3_substep_adjoint: q = three device addresses of the velocity the forward
        substep irk started from; ub, fb: DeviceFields holding the adjoints of the substep's
        output velocity and stored right-hand side (ub is overwritten); ubo, fbo receive those of
        its input velocity and the right-hand side it read; pp: scratch DeviceField.  With scalars
        (cans_rk3_substep_adjoint_scal) scal = (s, sb, fsb, sbo, fsbo): the device addresses of the
        scalars the substep started from, and per scalar the DeviceFields of the adjoints of its
        output and stored right-hand side (sb overwritten) and of those it read."""
        args = [C.c_void_p(x) for x in q] + [C.c_void_p(f.ptr) for f in ub + fb + ubo + fbo] + [C.c_void_p(pp.ptr)]
        if not scal:
            LB.check(LB.lib().cans_rk3_substep_adjoint(self.ctx.h, C.byref(self.prm), irk, dt, *args),
                     "cans_rk3_substep_adjoint")
            return
        n = len(self.s)
        s_addr, sb, fsb, sbo, fsbo = scal
        arrs = [(C.c_void_p * n)(*[C.c_void_p(x) for x in s_addr])]
        for fl in (sb, fsb, sbo, fsbo):
            arrs.append((C.c_void_p * n)(*[C.c_void_p(f.ptr) for f in fl]))
        LB.check(LB.lib().cans_rk3_substep_adjoint_scal(self.ctx.h, C.byref(self.prm), irk, dt, *args,
                                                        *[C.cast(a, C.POINTER(C.c_void_p)) for a in arrs]),
                 "cans_rk3_substep_adjoint_scal")

    def capture_step(self, dt, nsteps=1):
        """Record nsteps fused RK3 steps of size dt into a HIP graph (torch.cuda.CUDAGraph)
        on the context's own stream; replay() then repeats them without host launch cost.

        The context must have been created with its own stream, and one eager step must
        have run first (it allocates the scratch buffers and zeroes RHS_old; the fused
        step itself allocates, copies and synchronises nothing).  dt is baked into the
        graph; a new dt needs a new capture."""
        torch = self.ctx.torch
        if self.ctx.torch_stream == torch.cuda.default_stream(self.ctx.device):
            raise ValueError("capture_step needs a context created with its own stream")
        torch.cuda.synchronize()
        g = torch.cuda.CUDAGraph()
        with torch.cuda.graph(g, stream=self.ctx.torch_stream):
            for _ in range(nsteps):
                self.rk3_step(dt)
        return g

    def rk3_step_unfused(self, dt):
        """The same substeps through the per-routine entry points (main.f90:452-472)."""
        for irk in (1, 2, 3):
            rkpar = RKCOEFF[irk - 1]
            dtrk = (rkpar[0] + rkpar[1]) * dt
            dtrki = dtrk ** (-1)
            for i in range(len(self.s)):  # main.f90:440-450
                fs = self.rk_scal(i, irk, dt)
                self.bulk_forcing_s(i, fs)
                if self.imp:
                    self.splans[i].solve_helmholtz(self.s[i], -0.5 * self.scal_spec[i]["alpha"] * dtrk)
                self.boundp_s(i)
            self.rk(irk, dt, with_f=False)
            alpha = 0.0
            if self.imp:  # main.f90:453-462
                alpha = -.5 * self.grid.visc * dtrk
                for f, hp in zip((self.u, self.v, self.w), self.hplans):
                    hp.solve_helmholtz(f, alpha)
            self.bounduvw(self.u, self.v, self.w, 0, LB.HALO_LOWER)
            self.fillps(dtrki)
            self.solver()
            self.boundp(self.pp)
            self.correc(dtrk)
            self.bounduvw(self.u, self.v, self.w, 1, LB.HALO_UPPER, LB.HALO_NONE)
            self.updatep(alpha)
            self.boundp(self.p, LB.HALO_UPPER)

    def close(self):
        self.plan.close()
        for pl in self.hplans + [x for x in self.splans if x is not None]:
            pl.close()
        self.ctx.sync()
